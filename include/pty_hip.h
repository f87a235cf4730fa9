/*
 * pty_hip.h -- C ABI of the MI355X (gfx950) ptychography hot path.
 *
 * The reference (PtyLab 0.3.1) has no FFI: its hot path is the Python engine
 * API Engines.ePIE|mPIE|multiPIE|e3PIE(reconstruction, experimentalData,
 * params, monitor).reconstruct() (Engines/ePIE.py:20-51, mPIE.py:29-194,
 * multiPIE.py:29-123, e3PIE.py:29-190) and the propagator pair
 * Operators.object2detector/detector2object (Operators/Operators.py:561-584).
 * The entry points below are what a binding of that API needs; each comment
 * names the reference interface it replaces.  All pointers are device
 * pointers unless stated; arrays are C-order and contiguous; complex numbers
 * are interleaved float pairs (complex64).  Every function returns 0 on
 * success or a negative PTY_E* code, with a message in pty_last_error().
 * Nothing here synchronises the host unless stated.
 */
#ifndef PTY_HIP_H
#define PTY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTY_ABI_VERSION 4

/* error codes */
#define PTY_OK 0
#define PTY_EINVAL (-22)      /* bad argument / unsupported shape           */
#define PTY_ENOMEM (-12)      /* device allocation failed                   */
#define PTY_EDEVICE (-5)      /* HIP runtime error (launch, copy, ...)      */
#define PTY_ETIMEDOUT (-110)  /* a cluster barrier exceeded its spin limit  */

/* engines (Engines/ePIE.py, mPIE.py, multiPIE.py, e3PIE.py, qNewton.py) */
#define PTY_EPIE 0
#define PTY_MPIE 1
#define PTY_MULTIPIE 2
#define PTY_E3PIE 3
#define PTY_QNEWTON 4  /* Engines/qNewton.py; alphaObject/alphaProbe carry regObject/regProbe */
#define PTY_MQNEWTON 5 /* Engines/mqNewton.py with ADAM momentum: as PTY_QNEWTON, plus momFriction = beta1,
                          momFeedback = beta2, probeWeight = betaObject_m, sliceBeta = betaProbe_m; needs the
                          buffer and momentum arrays; second moments start at zero */

/* propagators (Operators.py:1082-1103 dispatch keys) */
#define PTY_FRAUNHOFER 0          /* fft2c with shifts, fftshiftSwitch=False  */
#define PTY_FRAUNHOFER_SHIFTED 1  /* fft2 ortho, fftshiftSwitch=True          */
#define PTY_ASP 2                 /* band-limited angular spectrum            */
#define PTY_FRESNEL 3             /* fft2c(psi * Q), fftshiftSwitch=False (Operators.py:77-152) */
#define PTY_POLYCHROME_ASP 4      /* ASP with one transfer function per wavelength (Operators.py:458-558) */
#define PTY_SCALED_ASP 5          /* ifft2c(fft2c(psi Q1) Q2), scaled ASP (Operators.py:313-370) */
#define PTY_SCALED_POLYCHROME_ASP 6 /* scaled ASP, Q1 and Q2 per wavelength (Operators.py:378-455) */

/*
 * One reconstruction.  Shapes follow the reference's 6D layout
 * (Reconstruction.py:356-413):
 *   object  (nlambda, nosm, 1, nslice, No, No)       complex64
 *   probe   (nlambda, 1, npsm, nslice, Np, Np)       complex64
 *   ptychogram (F, Np, Np) float32, natural detector order
 *   positions (F, 2) int32 [row, col] of the patch's upper-left pixel
 *             (Reconstruction.positions, Reconstruction.py:675-710)
 *   order   (n_iter, F) int32: position index visited at step j of iteration i
 *             (BaseEngine.setPositionOrder, BaseEngine.py:637-665)
 *   momentum_mask (n_iter, F) uint8: 1 where np.random.rand(1) > 0.95 fired
 *             after step j (mPIE.py:170-172); ignored by ePIE/e3PIE
 *   energy  (F) float32 energyAtPos (ExperimentalData.py:248)
 *   tf      (Np, Np) complex64 transfer function in FFT frequency order
 *             (ifftshift applied): the ASP detector TF (Operators.py:861-893)
 *             or the e3PIE slice TF (e3PIE.py:61-69); for PTY_POLYCHROME_ASP
 *             (nlambda, Np, Np), one per wavelength of spectralDensity
 *             (Operators.py:910-952); for PTY_FRESNEL the quadratic phase Q
 *             (Np, Np) in natural order (_propagation_kernels.py:16-36);
 *             for PTY_SCALED_ASP (2, Np, Np): Q2 as built on its centred
 *             grid (used unshifted) and Q1 (-1)^(y+x) in natural order
 *             (Operators.py:706-760, 955-987; nlambda == 1); for
 *             PTY_SCALED_POLYCHROME_ASP (2 nlambda, Np, Np): the nlambda Q2,
 *             then the nlambda Q1 (-1)^(y+x) (Operators.py:990-1032);
 *             NULL for Fraunhofer
 *   err_at_pos (n_iter, F) float64 and err_iter (n_iter) float64 outputs
 *             (getErrorMetrics, BaseEngine.py:819-841)
 * Buffers and momenta (mPIE/multiPIE only) have the object/probe shapes.
 */
typedef struct pty_instance {
  int32_t engine;
  int32_t propagator;
  int32_t nlambda, nosm, npsm, nslice;
  int32_t Np, No, F;
  int32_t n_iter;
  float betaObject, betaProbe;
  float alphaObject, alphaProbe;
  float momFriction;   /* mPIE frictionM / multiPIE stepM   */
  float momFeedback;   /* mPIE feedbackM / multiPIE betaM   */
  float probeWeight;   /* mPIE probe-update weight (1)      */
  float sliceBeta;     /* e3PIE inner-slice probe step (0.9, e3PIE.py:140) */
  void* object;
  void* probe;
  void* objectBuffer;
  void* objectMomentum;
  void* probeBuffer;
  void* probeMomentum;
  const float* ptychogram;
  const int32_t* positions;
  const int32_t* order;
  const uint8_t* momentum_mask;
  const float* energy;
  const void* tf;
  double* err_at_pos;
  double* err_iter;
  int32_t workgroups;  /* workgroups cooperating on this instance; 0 = auto */
  int32_t flags;       /* PTY_FLAG_* bits, 0 by default */
  /* Position correction (BaseEngine.positionCorrection, BaseEngine.py:1241-1323;
   * mPIE.py:163-166, pcPIE.py:107-108): with pc_direction set (F, 2) float64
   * search directions D, in/out, the PIE engines correlate every position's
   * pre-update patch, shifted by one pixel in the 9 directions, with the
   * updated patch and set D[k] = 0.5 clip(grad, 3) + 0.9 D[k], from
   * iteration pc_start of the plan on.  Needs nlambda = nosm = nslice = 1.
   * The positions are read at every position, so the caller may rewrite
   * them between launches (positionCorrectionUpdate, BaseEngine.py:1400-1438);
   * with pc_direction set they are checked against the object at every
   * position.  Instances with pc_direction never join the batched launch. */
  double* pc_direction;
  int32_t pc_start;
  /* params.positionCorrectionSwitch_radius: below 2 the direct correlation
   * above; from 2 on the Fourier branch (BaseEngine.py:1278-1313): the
   * cross-correlation of the mean-free patches over shifts -r..r in both
   * axes, betaGrad 5, clipped at 10 px. */
  int32_t pc_radius;
  /* Optional (F, Np, Np) float32 output, natural detector order: the last
   * iteration's |I_k - Iest| of every position k, the reference's
   * reconstruction.detectorError (getRMSD, BaseEngine.py:851-868; the
   * reference rewrites it at every iteration, so after a run it holds the
   * last iteration's values).  NULL: not written. */
  float* detector_error;
} pty_instance;

/* Run the instance in the batched launch (the one kernel that serves every
 * Np <= 256 of a pass family) even when it is alone in its plan.  In a plan
 * of several instances every instance with Np <= 256 runs there anyway. */
#define PTY_FLAG_BATCHED 1
#define PTY_FLAG_MQN_NADAM 2    /* PTY_MQNEWTON with momentum_method "NADAM" (mqNewton.py:155-174) */
#define PTY_FLAG_MQN_MOMENTUM 4 /* PTY_MQNEWTON with momentum_method "momentum" (mqNewton.py:176-182):
                                   momFriction carries frictionM, momFeedback is unused */

typedef struct pty_plan pty_plan;

/* ABI version of the loaded library (compare with PTY_ABI_VERSION). */
int pty_abi_version(void);

/* Message of the last failing call on this host thread. */
const char* pty_last_error(void);

/*
 * Build a plan for `ninst` independent reconstructions run by one persistent
 * launch per iteration.  Several instances of uneven Np and position counts
 * share one launch per pass family (the batched kernel); an instance's
 * `workgroups` sets its cluster size (choose it with the autotuner, see
 * pty_coresident_workgroups).  There is no batching API in the reference:
 * a batch is what a script looping over independent engine.reconstruct()
 * calls (ePIE.py:53-112, mPIE.py:111-180) would do one after another.
 * Replaces the set-up half of engine.reconstruct():
 * BaseEngine._prepareReconstruction / _checkGPU (BaseEngine.py:96-111,
 * 613-635) and the engines' initializeReconstructionParams.  Validates every
 * shape, allocates the per-instance workspace, permutes each ptychogram and
 * transfer function into the kernel's bit-reversed layout, uploads the
 * descriptors.  In the batched launch every cluster runs whole on one XCD:
 * an instance's `workgroups` may not exceed one XCD's co-resident share
 * (pty_coresident_workgroups / 8), and the clusters must pack into the 8
 * XCDs (largest first, each to the least-loaded XCD); PTY_EINVAL otherwise.
 * Synchronises the device once (set-up only).
 */
int pty_plan_create(pty_plan** plan, const pty_instance* insts, int32_t ninst, void* hip_stream);

/*
 * Run iterations [it_begin, it_end) of every instance, stream-ordered, no
 * host synchronisation, no allocation.  Replaces the body of
 * engine.reconstruct(): the per-position loop plus getErrorMetrics
 * (ePIE.py:69-112, mPIE.py:111-180, multiPIE.py:74-114, e3PIE.py:89-180).
 */
int pty_plan_run(pty_plan* plan, int32_t it_begin, int32_t it_end, void* hip_stream);

/*
 * Run positions [j_begin, j_end) of iteration `it` (every instance, clamped to
 * its frame count): the granularity of ePIE.reconstruct_stepwise, which
 * yields after every scan position (ePIE.py:53-112).  The iteration's error
 * is reduced when its last position has run.
 */
int pty_plan_run_positions(pty_plan* plan, int32_t it, int32_t j_begin, int32_t j_end, void* hip_stream);

/*
 * Read the plan's device-side failure flags; synchronises the stream.
 * Returns PTY_ETIMEDOUT if a cluster barrier gave up, PTY_EDEVICE if a
 * workgroup of the batched launch found no slot on its XCD, PTY_EINVAL if a
 * position of an instance with position correction, rewritten between
 * launches, put its patch outside the object (that instance stopped at it).
 */
int pty_plan_check(pty_plan* plan, void* hip_stream);

/*
 * Run every later pty_plan_run / pty_plan_run_positions of this plan as one
 * replay per iteration of an instantiated HIP graph (control-block reset
 * plus the class kernels, with the iteration arguments rewritten per
 * replay) instead of direct launches.  On by default for plans with a
 * batched launch.  No reference counterpart (a launch-mode switch).
 */
int pty_plan_use_graph(pty_plan* plan, int32_t enable);

/*
 * Workgroups of the kernel that would run an instance of this probe size,
 * engine and propagator (batched != 0: in the batched launch) that the
 * current device can hold at once: the upper bound for the sum of the
 * instances' `workgroups` in one launch.  Host query for the autotuner; no
 * reference counterpart.
 */
int pty_coresident_workgroups(int32_t Np, int32_t engine, int32_t propagator, int32_t batched, int32_t* count);

/* Number of workgroups that work per iteration (host query; blocks left
 * idle by the XCD packing of the batched launch are not counted). */
int pty_plan_workgroups(const pty_plan* plan, int32_t* total);

/* Release a plan (device-synchronising). */
int pty_plan_destroy(pty_plan* plan);

/*
 * In-place propagation of `nfields` contiguous (Np, Np) complex64 fields:
 * Operators.object2detector (inverse=0) / detector2object (inverse=1) for
 * the Fraunhofer and ASP propagators (Operators.py:27-74, 155-230).  `tf`
 * is the ASP transfer function in FFT frequency order (NULL for Fraunhofer).
 */
int pty_propagate(void* fields, int32_t nfields, int32_t Np, int32_t propagator,
                  int32_t inverse, const void* tf, void* hip_stream);

/*
 * Plain batched 2D DFT of `nfields` (Np, Np) complex64 fields, in place:
 * sign=-1 forward, +1 inverse, unnormalised (numpy.fft.fft2 / ifft2 * N^2).
 * The primitive behind utils.fft2c/ifft2c (utils.py:12-47).
 */
int pty_fft2(void* fields, int32_t nfields, int32_t Np, int32_t sign, void* hip_stream);

/*
 * Per-iteration constraints of BaseEngine.applyConstraints (BaseEngine.py:
 * 1440-1612) on device state, between pty_plan_run calls.
 */

/* out[0] = sum of |x|^2 over n complex64 values, FP64, deterministic: the
 * probe power of the power correction (BaseEngine.py:1462-1469, 227-235). */
int pty_field_power(const void* x, size_t n, double* out, void* hip_stream);

/* x *= target / sqrt(power[0]) (power read on the device): the probe power
 * correction to ExperimentalData.maxProbePower (ExperimentalData.py:250). */
int pty_scale_to_power(void* x, size_t n, const double* power, float target, void* hip_stream);

/* x = (1 - aleph) x + aleph x w over nfields (N, N) fields, w an (N, N)
 * float32 window: the absorbing probe boundary (aleph < 1,
 * BaseEngine.py:1496-1506) and the probe boundary (aleph = 1, x *= w,
 * :1493-1494) with the windows of _probeWindow (:237-266). */
int pty_window_fields(void* x, int32_t nfields, int32_t N, const float* window, float aleph, void* hip_stream);

/* out3 = (sum |x| X, sum |x| Y, sum |x|) over one (N, N) field, FP64: the
 * centre of mass of comStabilization (BaseEngine.py:1724-1745). */
int pty_field_moments(const void* x, int32_t N, const float* X, const float* Y, double* out3, void* hip_stream);

/* numpy.roll of nfields (Ny, Nx) fields by (dy, dx) in place, through a
 * scratch buffer of the same size: the probe/object shift of
 * comStabilization (BaseEngine.py:1747-1776). */
int pty_roll_fields(void* x, void* scratch, int32_t nfields, int32_t Ny, int32_t Nx, int32_t dy, int32_t dx,
                    void* hip_stream);

/* pty_roll_fields without a workspace and without any allocation: no
 * counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * scratch is a static device array of PTY_ROLL_SCRATCH_MAX_VALUES complex64
 * values (16 MiB) per GPU, held until the process ends; nfields*Ny*Nx above
 * that returns PTY_EINVAL.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (different streams or host threads) race.
 * Two kernel launches, no memset or memcpy nodes. */
#define PTY_ROLL_SCRATCH_MAX_VALUES 2097152 /* 2 Mi complex64 values = 16 MiB */
int pty_roll_fields_noalloc(void* x, int32_t nfields, int32_t Ny, int32_t Nx, int32_t dy, int32_t dx,
                            void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* PTY_HIP_H */
