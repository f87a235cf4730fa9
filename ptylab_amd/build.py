"""Build ``libptyhip.so`` in-tree with hipcc for gfx950."""

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
REPO = HERE.parent
SRC = [HERE / "csrc" / "pty_hip.hip", HERE / "csrc" / "pty_constraints.hip"]
OUT = HERE / "libptyhip.so"
NET_SRC = [HERE / "csrc" / "pty_net.hip"]
NET_OUT = HERE / "libptynet.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    # the reference computes every product and sum of the update rules
    # unfused (NumPy elementwise ops); keep the same rounding
    "-ffp-contract=off",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    # packed FP32 instructions stay off
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# the position-loop kernels are latency-bound (one item per wave per pass):
# the instruction scheduler's ILP strategy keeps more loads in flight.  On
# one box: mPIE ASP Np = 256 (the benchmark) 27.09 -> 26.65 us per position,
# ePIE Fraunhofer Np = 64 13.43 -> 13.61, e3PIE Np = 128 72.33 -> 72.67, the
# batched launches unchanged within their spread
HOT_PATH_FLAGS = ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]


def _compile(srcs, out, verbose, extra=()):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc, *HIPCC_FLAGS, *extra, f"-I{REPO / 'include'}", "-o", str(tmp), *[str(s) for s in srcs]]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {out.name} ({res.returncode}):\n{res.stderr[-4000:]}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


def build(verbose=False):
    """Compile both libraries: the reconstruction hot path (libptyhip.so)
    and the learned initialiser's network kernels (libptynet.so)."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(2) as ex:
        f1 = ex.submit(_compile, SRC, OUT, verbose, HOT_PATH_FLAGS)
        f2 = ex.submit(_compile, NET_SRC, NET_OUT, verbose)
        f2.result()
        return f1.result()


if __name__ == "__main__":
    build(verbose=True)
