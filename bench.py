"""Benchmark: position updates/s of the mPIE reconstruction loop on one MI355X.

Workload (configuration 1 of the project's benchmark set): mPIE with the
band-limited angular-spectrum propagator, complex64, wavelength 632.8 nm,
dxo 4 um, z 50 mm, object No=1024 (smoothed random amplitude/phase), probe
Np=256 with npsm=4 orthogonalised Hermite-Gaussian x aperture modes, 1000
jittered Fermat-spiral positions, 1e5 photons peak with Poisson noise,
frictionM 0.7 / feedbackM 0.3.  Synthetic data (no network).

A step is one iteration: all 1000 positions in the engine's order, one
persistent launch.  W warm-up steps, then K timed steps between
synchronisations; value = K * F / elapsed.  The inputs are resident in HBM
before the timed region.

Extra fields: ``roofline`` for the persistent kernel (algorithmic bytes per
launch / launch time from HIP events on the launch stream, against 8 TB/s),
``cpu_baseline`` (the NumPy oracle timed on one host core over three full
iterations of the same workload, about 13 s).
"""

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "position updates/s; time to object NRMSE 1e-2, learned-init vs plain mPIE"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 (vector FMA) peak
TRAFFIC_FILE = ROOT / "profiles" / "pty_iterate_kernel_mpie_asp_np256_traffic.txt"
BATCH_TRAFFIC_FILE = ROOT / "profiles" / "pty_batch_kernel_mpie_asp_64inst_traffic.txt"
BATCH_EPIE_TRAFFIC_FILE = ROOT / "profiles" / "pty_batch_kernel_epie_fraunhofer_64inst_traffic.txt"


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # 40 timed iterations: about a second of GPU time for the default
    # workload, long enough for an outside utilisation sampler to see it
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-positions", type=int, default=3000,
                    help="position updates of the CPU baseline sample (whole iterations of the workload when "
                         "larger than its position count)")
    ap.add_argument("--workgroups", type=int, default=0)
    ap.add_argument("--F", type=int, default=1000)
    ap.add_argument("--workload", choices=("c1", "c0", "c2", "batched", "learned_init"), default="c1",
                    help="c1: single mPIE ASP reconstruction (default); c0: single plain ePIE Fraunhofer "
                         "reconstruction (Np=64, 100 positions; microseconds per position); c2: single multislice "
                         "reconstruction (four slices, Np=128, 400 positions); batched: 64 independent "
                         "reconstructions; learned_init: time to object NRMSE, learned-init vs plain mPIE")
    ap.add_argument("--c2-engine", choices=("e3PIE", "multiPIE"), default="e3PIE",
                    help="c2: the multislice engine (e3PIE: the slice chain; multiPIE: slices as independent axes)")
    ap.add_argument("--no-orthogonalization", action="store_true",
                    help="c1: skip the probe-mode orthogonalisation every 10 iterations")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="batched: processes of the all-cores CPU baseline (one instance each)")
    ap.add_argument("--train-patterns", type=int, default=200000)
    ap.add_argument("--li-stride", type=int, default=4,
                    help="learned_init: predict and stitch the patches of every n-th position (the fallback check "
                         "uses all of them)")
    ap.add_argument("--li-known-probe", action="store_true",
                    help="learned_init: both sides start mPIE from the known probe (the one the network's features "
                         "use) instead of the reference's default 'circ' guess")
    ap.add_argument("--engine", choices=("mPIE", "ePIE"), default="mPIE", help="engine of the batched workload")
    ap.add_argument("--instances", type=int, default=64)
    ap.add_argument("--gen-workers", type=int, default=16,
                    help="batched: processes generating the inputs (1: in this process, no fork; for runs under a "
                         "profiler, which initialises the GPU before the program starts)")
    ap.add_argument("--no-graph", action="store_true", help="batched: direct launches instead of graph replays")
    ap.add_argument("--batched-inputs", choices=("asp_modes", "fraunhofer"), default="asp_modes",
                    help="batched: instances generated as the single mPIE ASP workload (ASP, 4 probe modes; "
                         "default) or as the plain-ePIE family (Fraunhofer, one mode)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy before the timed region (profiler counter scale)")
    return ap.parse_args()


def bytes_per_position(Np, No, nl, no, npm, ns, momentum_fraction):
    """Algorithmic HBM bytes of one position update: object patch and probe
    read and written once (complex64), the measured frame read once, plus the
    momentum step's 6 full-array passes at the triggered fraction."""
    core = 16 * Np * Np * (nl * no * ns + nl * npm * ns) + 4 * Np * Np
    mom = 48 * (nl * no * ns * No * No + nl * npm * ns * Np * Np) * momentum_fraction
    return core + mom


def fft_flops_per_position(Np, n_fft):
    """Nominal FFT work of one position update: n_fft 2D transforms of
    5 Np^2 log2(Np^2) flops each (the radix-2 operation count)."""
    return n_fft * 5.0 * Np * Np * np.log2(float(Np * Np))


def host_info():
    """CPU model, usable cores and NumPy version of the host that ran the
    CPU baseline."""
    model = None
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "host_cores": usable, "numpy": np.__version__}


def read_traffic(path=TRAFFIC_FILE):
    if not path.exists():
        return None
    txt = path.read_text()
    m = re.search(r"traffic_bytes_per_launch\s*=\s*([0-9.eE+]+)", txt)
    return float(m.group(1)) if m else None


def batch_traffic(args):
    """Measured HBM bytes per launch of the batched workload, where a
    profile of that exact workload is committed."""
    if args.instances != 64:
        return None
    if args.engine == "mPIE" and args.batched_inputs == "asp_modes":
        return read_traffic(BATCH_TRAFFIC_FILE)
    if args.engine == "ePIE" and args.batched_inputs == "fraunhofer":
        return read_traffic(BATCH_EPIE_TRAFFIC_FILE)
    return None


def cpu_baseline(workload, n_positions):
    import oracle as O

    w = workload
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No, npsm=4)
    np.random.seed(1)
    obj = O.initial_guess(geo.shape_O, "ones", geo)
    pr = O.initial_guess(geo.shape_P, "circ", geo)
    F = w.ptychogram.shape[0]
    iters = max(1, -(-n_positions // F))
    per_iter = min(n_positions, F)
    run = O.OracleRun(geo, obj, pr, engine="mPIE", propagator="ASP", numIterations=iters)
    t0 = time.perf_counter()
    run.reconstruct(positions_limit=per_iter)
    dt = time.perf_counter() - t0
    n = iters * per_iter
    return {
        "value": n / dt,
        "unit": "position updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"NumPy oracle, mPIE ASP npsm=4 Np=256 No=1024: {iters} iteration(s) of {per_iter} positions "
                  f"({n} updates, {dt:.1f} s)",
        **host_info(),
    }


def main():
    args = parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); independent reconstructions "
              "shard across GPUs without communication, which this benchmark does not measure", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        return main_batched(args)
    if args.workload == "learned_init":
        return main_learned_init(args)
    if args.workload == "c0":
        return main_c0(args)
    if args.workload == "c2":
        return main_c2(args)
    import torch

    import ptylab_amd as P
    from ptylab_amd import synthetic as S

    if not torch.cuda.is_available():
        print("no HIP GPU available", file=sys.stderr)
        sys.exit(1)
    t_gen = time.perf_counter()
    w = S.cpm_asp_modes(Np=256, No=1024, F=args.F, npsm=4, seed=10)
    print(f"[bench] workload generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)

    data = P.ExperimentalData()
    for k, v in w.as_dict().items():
        setattr(data, k, v)
    data._setData()
    params = P.Params()
    params.propagatorType = "ASP"
    params.positionOrder = "random"
    # "feedback every 10 iterations": the snapShots orthogonalisation of the
    # four probe modes every 10 iterations (Params.py:78-79)
    params.orthogonalizationSwitch = not args.no_orthogonalization
    params.orthogonalizationFrequency = 10
    rec = P.Reconstruction(data, params)
    rec.No = w.No
    rec.npsm = 4
    np.random.seed(1)
    rec.initializeObjectProbe()
    eng = P.Engines.mPIE(rec, data, params, P.DummyMonitor())
    eng.frictionM, eng.feedbackM = 0.7, 0.3
    eng.workgroups = args.workgroups
    K, W = args.steps, args.warmup
    eng.numIterations = W + K
    np.random.seed(1)
    eng._prepareReconstruction()
    plan, T = eng._begin()
    stream = torch.cuda.current_stream()
    F = data.numFrames
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    orth = bool(params.orthogonalizationSwitch)
    for it in range(W):
        plan.run(it, it + 1)
        if orth:
            eng.applyConstraints(it)
    plan.check()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    t0 = time.perf_counter()
    for i in range(K):
        ev[i][0].record(stream)
        plan.run(W + i, W + i + 1)
        ev[i][1].record(stream)
        if orth:
            # the probe modes' snapShots orthogonalisation at every 10th
            # iteration (Params.orthogonalizationFrequency), inside the clock
            eng.applyConstraints(W + i)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    plan.check()
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    masks = eng._masks[W:W + K]
    n_wg = plan.workgroups()
    eng._end(W + K)
    triggers = float(masks.sum()) / (K * F)
    bpos = bytes_per_position(256, w.No, 1, 1, 4, 1, triggers)
    mean_launch_s = float(np.mean(launch_ms)) / 1e3
    achieved = bpos * F / mean_launch_s / 1e9
    traffic = read_traffic()
    fpos = fft_flops_per_position(256, 4 * 4)  # ASP: 4 transforms per mode and position, 4 modes
    out = {
        "metric": METRIC,
        "value": K * F / elapsed,
        "unit": "position updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": "mPIE, ASP, npsm=4 modes, Np=256, No=1024, F=%d positions per step (configuration 1)" % F,
            "positions_per_step": F,
            "engine_workgroups": n_wg,
            "orthogonalization_every": 10 if orth else None,
            "us_per_position": mean_launch_s / F * 1e6,
        },
        "fft": {
            "flops_per_position": fpos,
            "achieved": fpos * F / mean_launch_s / 1e12,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": fpos * F / mean_launch_s / 1e12 / FP32_PEAK_TFLOPS,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "pty_iterate_kernel",
            "bytes_per_position": bpos,
            "mean_launch_ms": mean_launch_s * 1e3,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_positions)
    print(json.dumps(out))


def cpu_baseline_batched(workloads, engine, propagator, npsm, budget_s=15.0):
    """The NumPy oracle on one host core over the batch's instances in order,
    one iteration each (at most the positions left in the budget's estimate),
    until about budget_s seconds of work."""
    import oracle as O

    done, t_all = 0, 0.0
    names = []
    for w in workloads:
        geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No,
                         npsm=npsm)
        np.random.seed(1)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        run = O.OracleRun(geo, obj, pr, engine=engine, propagator=propagator, numIterations=1)
        t0 = time.perf_counter()
        run.reconstruct()
        t_all += time.perf_counter() - t0
        done += w.ptychogram.shape[0]
        names.append(f"Np={w.Np}/F={w.ptychogram.shape[0]}")
        if t_all >= budget_s:
            break
    return {
        "value": done / t_all,
        "unit": "position updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"NumPy oracle, {engine} {propagator} npsm={npsm}, one iteration of the first {len(names)} "
                  f"batch instances "
                  f"({', '.join(names)}; {done} positions, {t_all:.1f} s)",
        **host_info(),
    }


def _oracle_instance_sample(args):
    """Oracle iterations of one batch instance in a worker process, in chunks
    of at most `chunk` positions, until `budget` seconds of work; returns
    (positions, seconds) measured inside the worker."""
    import oracle as O

    w, engine, propagator, npsm, chunk, budget = args
    geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No,
                     npsm=npsm)
    np.random.seed(1)
    obj = O.initial_guess(geo.shape_O, "ones", geo)
    pr = O.initial_guess(geo.shape_P, "circ", geo)
    run = O.OracleRun(geo, obj, pr, engine=engine, propagator=propagator, numIterations=1)
    n = min(chunk, w.ptychogram.shape[0])
    done, t = 0, 0.0
    while t < budget:
        t0 = time.perf_counter()
        run.reconstruct(positions_limit=n)
        t += time.perf_counter() - t0
        done += n
    return done, t


def cpu_baseline_all_cores(workloads, engine, propagator, npsm, procs, chunk=100, budget=8.0):
    """All-cores CPU throughput of the batch: `procs` forked processes, one
    instance each (the first `procs` instances), each running oracle
    positions for about `budget` seconds, which gives every sampled
    instance's seconds per position c_i with all cores busy (measured inside
    the workers, so process start-up is excluded).  One iteration of the
    sampled instances, spread evenly over the cores, takes sum F_i c_i /
    procs, so the aggregate is procs * sum F_i / sum F_i c_i (a sum of the
    per-process rates would overweight the cheap small-Np instances).  Runs
    before this process initialises the GPU (fork)."""
    import multiprocessing as mp

    procs = max(1, min(procs, len(workloads)))
    jobs = [(w, engine, propagator, npsm, chunk, budget) for w in workloads[:procs]]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(procs) as pool:
        res = pool.map(_oracle_instance_sample, jobs, chunksize=1)
    wall = time.perf_counter() - t0
    n = int(sum(r[0] for r in res))
    Fs = [w.ptychogram.shape[0] for w in workloads[:procs]]
    work = sum(F * r[1] / r[0] for F, r in zip(Fs, res))  # seconds of one iteration on one core
    return {
        "value": float(procs * sum(Fs) / work),
        "unit": "position updates/s",
        "cores": procs,
        "kind": "port",
        "sample": f"NumPy oracle, {engine} {propagator} npsm={npsm}: {procs} concurrent processes, one batch instance "
                  f"each (the first {procs}), positions in chunks of at most {chunk} for about {budget:.0f} s each "
                  f"({n} updates, {wall:.1f} s wall); value = cores x positions of one iteration of the sampled "
                  f"instances / its single-core seconds at the measured per-position costs",
        **host_info(),
    }


def main_batched(args):
    """Configuration "batched": independent mPIE (or ePIE) reconstructions,
    Np uniform in {64, 128, 256}, No = 4 Np, F uniform in [100, 800], inputs
    generated as in the single mPIE ASP workload (ASP, 4 probe modes, 1e5
    photons, per-instance seeds; --batched-inputs fraunhofer for the
    plain-ePIE family), one launch per iteration replayed as a HIP graph,
    cluster sizes from the autotuner.  A step is one iteration of every
    instance; value = K * sum(F) / elapsed."""
    import torch

    import ptylab_amd as P
    from ptylab_amd import autotune
    from ptylab_amd import synthetic as S

    # inputs first: the generator forks worker processes, which must happen
    # before this process initialises the GPU
    t_gen = time.perf_counter()
    ws = S.batched_instances(n=args.instances, workers=min(args.gen_workers, os.cpu_count() or 1),
                             inputs=args.batched_inputs,
                             progress=lambda i: print(f"[bench] input {i + 1}/{args.instances}", file=sys.stderr,
                                                      flush=True))
    prop, npsm = ("ASP", 4) if args.batched_inputs == "asp_modes" else ("Fraunhofer", 1)
    print(f"[bench] {len(ws)} workloads generated in {time.perf_counter() - t_gen:.1f} s", file=sys.stderr, flush=True)
    all_cores = None
    if not args.no_cpu_baseline and args.cpu_procs > 1:
        # forked processes: before this process initialises the GPU
        print("[bench] timing the all-cores CPU oracle baseline", file=sys.stderr, flush=True)
        all_cores = cpu_baseline_all_cores(ws, args.engine, prop, npsm, args.cpu_procs)
    if not torch.cuda.is_available():
        print("no HIP GPU available", file=sys.stderr)
        sys.exit(1)
    K, W = args.steps, args.warmup
    engines = []
    for i, w in enumerate(ws):
        data = P.ExperimentalData()
        for k, v in w.as_dict().items():
            setattr(data, k, v)
        data._setData()
        params = P.Params()
        params.propagatorType = prop
        params.positionOrder = "random"
        rec = P.Reconstruction(data, params)
        rec.No = w.No
        rec.npsm = npsm
        np.random.seed(1000 + i)
        rec.initializeObjectProbe()
        eng = getattr(P.Engines, args.engine)(rec, data, params, P.DummyMonitor())
        if args.engine == "mPIE":
            eng.frictionM, eng.feedbackM = 0.7, 0.3
        eng.numIterations = W + K
        engines.append(eng)
    np.random.seed(3)
    for e in engines:
        e._prepareReconstruction()
    staged = [e._stage() for e in engines]
    insts = [s_[0] for s_ in staged]
    for inst in insts:
        inst.flags |= P._native.PTY_FLAG_BATCHED
    t_tune = time.perf_counter()
    G = autotune.tune(insts, tensors=[s_[1] for s_ in staged])
    print(f"[bench] autotuned cluster sizes in {time.perf_counter() - t_tune:.1f} s: total {sum(G)} workgroups",
          file=sys.stderr, flush=True)
    for inst, g in zip(insts, G):
        inst.workgroups = g
    stream = torch.cuda.current_stream()
    plan = P._native.Plan(insts, [s_[1] for s_ in staged], stream)
    plan.use_graph(not args.no_graph)
    Fs = np.array([e.experimentalData.numFrames for e in engines])
    if args.calibrate_copy:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for it in range(W):
        plan.run(it, it + 1)
    plan.check()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    t0 = time.perf_counter()
    for i in range(K):
        ev[i][0].record(stream)
        plan.run(W + i, W + i + 1)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    plan.check()
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    n_wg = plan.workgroups()
    momentum = args.engine == "mPIE"
    bytes_iter = 0.0
    flops_iter = 0.0
    n_fft = (4 if prop == "ASP" else 2) * npsm
    for e, w in zip(engines, ws):
        frac = float(e._masks[W:W + K].sum()) / (K * e.experimentalData.numFrames) if momentum else 0.0
        bytes_iter += e.experimentalData.numFrames * bytes_per_position(w.Np, w.No, 1, 1, npsm, 1, frac)
        flops_iter += e.experimentalData.numFrames * fft_flops_per_position(w.Np, n_fft)
    for e in engines:
        e._collect(W + K)
    plan.close()
    mean_launch_s = float(np.mean(launch_ms)) / 1e3
    achieved = bytes_iter / mean_launch_s / 1e9
    nps = {int(n): int(sum(1 for w in ws if w.Np == n)) for n in sorted(set(w.Np for w in ws))}
    out = {
        "metric": METRIC,
        "value": K * float(Fs.sum()) / elapsed,
        "unit": "position updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": f"{len(ws)} independent {args.engine} reconstructions, {prop}, npsm={npsm}, Np in {nps}, No=4Np, "
                        f"F uniform in [100, 800] (sum {int(Fs.sum())}), one launch per iteration"
                        + (" (graph replay)" if not args.no_graph else " (direct launch)"),
            "positions_per_step": int(Fs.sum()),
            "workgroups": n_wg,
            "cluster_sizes": {str(n): sorted(set(int(g) for g, w in zip(G, ws) if w.Np == n)) for n in nps},
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": batch_traffic(args),
            "kernel": "pty_batch_kernel",
            "bytes_per_step": bytes_iter,
            "mean_launch_ms": mean_launch_s * 1e3,
        },
        "fft": {
            "flops_per_step": flops_iter,
            "achieved": flops_iter / mean_launch_s / 1e12,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": flops_iter / mean_launch_s / 1e12 / FP32_PEAK_TFLOPS,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_batched(ws, args.engine, prop, npsm)
        if all_cores is not None:
            out["cpu_baseline_all_cores"] = all_cores
    print(json.dumps(out))


def _single_run(args, w, engine, propagator, nslice, dz, bpos_fn, fpos, workload, oracle_iters, oracle_kw):
    """One reconstruction of workload ``w`` with the package's engine: W
    warm-up iterations, then K timed ones (one launch each, HIP events on the
    launch stream); ``bpos_fn(momentum_fraction)`` gives the algorithmic
    bytes of one position update.  The CPU baseline times the NumPy oracle
    on the same inputs for ``oracle_iters`` iterations."""
    import torch

    import ptylab_amd as P

    if not torch.cuda.is_available():
        print("no HIP GPU available", file=sys.stderr)
        sys.exit(1)
    data = P.ExperimentalData()
    for k, v in w.as_dict().items():
        setattr(data, k, v)
    data._setData()
    params = P.Params()
    params.propagatorType = propagator
    params.positionOrder = "random"
    rec = P.Reconstruction(data, params)
    rec.No = w.No
    rec.nslice = nslice
    if nslice > 1:
        rec.dz = dz
        rec.refrIndex = 1.0
    np.random.seed(3)
    rec.initializeObjectProbe()
    eng = getattr(P.Engines, engine)(rec, data, params, P.DummyMonitor())
    eng.workgroups = args.workgroups
    K, W = args.steps, args.warmup
    eng.numIterations = W + K
    np.random.seed(3)
    eng._prepareReconstruction()
    plan, T = eng._begin()
    stream = torch.cuda.current_stream()
    F = data.numFrames
    for it in range(W):
        plan.run(it, it + 1)
    plan.check()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(K)]
    t0 = time.perf_counter()
    for i in range(K):
        ev[i][0].record(stream)
        plan.run(W + i, W + i + 1)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    plan.check()
    launch_ms = [a.elapsed_time(b) for a, b in ev]
    n_wg = plan.workgroups()
    masks = getattr(eng, "_masks", None)
    frac = float(masks[W:W + K].sum()) / (K * F) if masks is not None and engine in ("mPIE", "multiPIE") else 0.0
    eng._end(W + K)
    bpos = bpos_fn(frac)
    mean_launch_s = float(np.mean(launch_ms)) / 1e3
    achieved = bpos * F / mean_launch_s / 1e9
    out = {
        "metric": METRIC,
        "value": K * F / elapsed,
        "unit": "position updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": workload,
            "positions_per_step": F,
            "engine_workgroups": n_wg,
            "us_per_position": mean_launch_s / F * 1e6,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "pty_iterate_kernel",
            "bytes_per_position": bpos,
            "mean_launch_ms": mean_launch_s * 1e3,
        },
        "fft": {
            "flops_per_position": fpos,
            "achieved": fpos * F / mean_launch_s / 1e12,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": fpos * F / mean_launch_s / 1e12 / FP32_PEAK_TFLOPS,
        },
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        import oracle as O

        geo = O.Geometry(w.ptychogram, w.encoder, w.wavelength, w.dxd, w.zo, w.entrancePupilDiameter, No=w.No,
                         nslice=nslice)
        np.random.seed(3)
        obj = O.initial_guess(geo.shape_O, "ones", geo)
        pr = O.initial_guess(geo.shape_P, "circ", geo)
        run = O.OracleRun(geo, obj, pr, engine=engine, propagator=propagator, numIterations=oracle_iters,
                          dz=dz, **oracle_kw)
        t0 = time.perf_counter()
        run.reconstruct()
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": oracle_iters * F / dt, "unit": "position updates/s", "cores": 1,
                               "kind": "port",
                               "sample": f"NumPy oracle, {engine} {propagator} Np={w.Np} No={w.No} nslice={nslice}: "
                                         f"{oracle_iters} iteration(s) of {F} positions ({oracle_iters * F} updates, "
                                         f"{dt:.1f} s)", **host_info()}
    print(json.dumps(out))


def main_c0(args):
    """Plain ePIE, Fraunhofer, one reconstruction (the CPU-runnable
    configuration: Np = 64, No = 256, 100 positions, 1e4 photons): a step is
    one iteration (one launch); reported as microseconds per position and
    as the (small) HBM fraction of a latency-bound single reconstruction."""
    from ptylab_amd import synthetic as S

    w = S.cpm_fraunhofer(Np=64, No=256, F=100, seed=0)
    _single_run(args, w, "ePIE", "Fraunhofer", 1, 0.0, lambda frac: bytes_per_position(64, w.No, 1, 1, 1, 1, 0.0),
                fft_flops_per_position(64, 2),
                "ePIE, Fraunhofer, Np=64, No=256, F=100 positions per step (configuration 0)", 20, {})


def main_c2(args):
    """Multislice configuration: four weak-phase slices 5 um apart (ASP
    between slices), Np = 128, No = 512, 400 positions, Fraunhofer to the
    detector, 1e4 photons with Poisson noise.  e3PIE runs the slice chain
    (3 forward and 3 backward slice propagations plus the detector pair per
    position: 14 transforms); multiPIE treats the slices as independent
    broadcast axes (8 transforms, momentum at the mPIE rate).  A step is one
    iteration of the 400 positions."""
    from ptylab_amd import synthetic as S

    Np, ns = 128, 4
    w = S.cpm_multislice(Np=Np, No=512, F=400, nslice=ns, dz=5e-6, seed=20)
    if args.c2_engine == "e3PIE":
        n_fft = 2 + 4 * (ns - 1)
        bfn = lambda frac: bytes_per_position(Np, w.No, 1, 1, 1, ns, 0.0)  # noqa: E731
    else:
        n_fft = 2 * ns
        bfn = lambda frac: bytes_per_position(Np, w.No, 1, 1, 1, ns, frac)  # noqa: E731
    _single_run(args, w, args.c2_engine, "Fraunhofer", ns, 5e-6, bfn, fft_flops_per_position(Np, n_fft),
                f"{args.c2_engine}, multislice nslice=4 dz=5um, Fraunhofer, Np=128, No=512, F=400 positions per step "
                "(configuration 2)", 2, {})


def main_learned_init(args):
    """Learned-initialiser workload: Np = 128, No = 512, 500 positions,
    Fraunhofer, 1e4 photons peak with Poisson noise, held-out object.  The
    network is trained first on simulated patterns (reported, not part of
    the race); then plain mPIE from the uniform object and mPIE from the
    learned object run one iteration at a time until the object NRMSE is at
    or below each threshold.  The learned side's clock includes inference,
    stitching and the fallback check."""
    import torch

    import ptylab_amd as P
    from ptylab_amd import learned_init as LI
    from ptylab_amd import unet as U

    if not torch.cuda.is_available():
        print("no HIP GPU available", file=sys.stderr)
        sys.exit(1)
    w = LI.held_out_scan()
    probe = LI.probe()
    net = U.UNet(64, seed=0)
    steps = max(1, args.train_patterns // net.B)
    t0 = time.perf_counter()
    losses = LI.train(net, steps, seed=100, lr=1e-3, lr_final=5e-5, log_every=max(1, steps // 5))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    true = torch.as_tensor(w.object_true[0, 0, 0, 0], device="cuda")
    roi = torch.as_tensor(LI.roi_mask(w.positions, probe, w.No), device="cuda")
    max_iter = 60
    thresholds = (1e-2, 0.12, 0.10, 0.09, 0.08, 0.075, 0.07)

    def race(learned):
        data = P.ExperimentalData()
        for k, v in w.as_dict().items():
            setattr(data, k, v)
        data._setData()
        params = P.Params()
        params.propagatorType = "Fraunhofer"
        params.positionOrder = "random"
        rec = P.Reconstruction(data, params)
        rec.No = w.No
        np.random.seed(1)
        rec.initializeObjectProbe()
        if args.li_known_probe:
            rec.probe = np.broadcast_to(probe.astype(np.complex64), rec.shape_P).copy()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = LI.initialize_object(rec, data, net, probe, stride=args.li_stride) if learned else None
        eng = P.Engines.mPIE(rec, data, params, P.DummyMonitor())
        eng.numIterations = max_iter
        np.random.seed(2)
        eng._prepareReconstruction()
        plan, T = eng._begin()
        curve, ts = [], []
        for it in range(max_iter):
            plan.run(it, it + 1)
            curve.append(float(LI.nrmse_gpu(T["object"][0, 0, 0, 0], true, roi)))
            ts.append(time.perf_counter() - t0)
        eng._end(max_iter)
        return np.array(curve), np.array(ts), info

    # one untimed run of each side first: plan kernels, NRMSE reductions, the
    # network's inference and the stitching kernels are then loaded alike
    race(False)
    race(True)
    cp, tp, _ = race(False)
    cl, tl, info = race(True)
    res = {}
    for th in thresholds:
        hp = int(np.argmax(cp <= th)) if np.any(cp <= th) else None
        hl = int(np.argmax(cl <= th)) if np.any(cl <= th) else None
        res[str(th)] = {
            "plain_s": None if hp is None else float(tp[hp]), "plain_iterations": None if hp is None else hp + 1,
            "learned_s": None if hl is None else float(tl[hl]), "learned_iterations": None if hl is None else hl + 1,
            "plain_over_learned": None if hp is None or hl is None else float(tp[hp] / tl[hl]),
        }
    target = res[str(1e-2)]
    out = {
        "metric": METRIC,
        "value": target["plain_over_learned"],
        "unit": "plain / learned wall time to object NRMSE <= 1e-2 (null: not reached by both)",
        "n_gpus": 1,
        "steps": max_iter,
        "warmup": 1,
        "ms_per_step": float(np.mean(np.diff(tp))) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 (reconstruction), bf16 MFMA / fp32 accumulate (network)",
        "data": "synthetic",
        "config": {
            "workload": "learned initialiser + mPIE: Np=128, No=512, 500 positions, Fraunhofer, 1e4 photons, "
                        "held-out object; network trained on %d simulated patterns" % (steps * net.B),
            "initial_probe": "known (the network's)" if args.li_known_probe else "circ (reference default)",
            "train_seconds": t_train,
            "train_final_loss": float(np.mean(losses[-50:])),
            "init": info,
        },
        "thresholds": res,
        "nrmse_min": {"plain": float(cp.min()), "learned": float(cl.min())},
        "nrmse_first_iterations": {"plain": [round(float(x), 4) for x in cp[:6]],
                                   "learned": [round(float(x), 4) for x in cl[:6]]},
        "roofline": None,
        "cpu_baseline": None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
