// MI355X (gfx950) kernels and C ABI for the per-position ptychography loop.
//
// One persistent launch runs whole iterations of one or many independent
// reconstructions ("instances").  Instance i owns G_i workgroups (a cluster);
// positions of one instance are strictly sequential (position k+1 gathers the
// object that position k wrote, ePIE.py:84-105), so the parallelism is the
// pixels x modes x slices of one position and the instances of a batch.
//
// Per position the cluster alternates row passes and column passes over the
// (fields, Np, Np) exit-wave workspace W; a wave owns one line (a row or a
// column) of every field at a time and keeps all fields of that line in
// registers, so the mode sums of the intensity constraint and of the
// object/probe updates are register-local.  Cluster-wide ordering between
// passes is an agent-scope release/acquire barrier on a per-instance
// counter; with G == 1 it is a plain workgroup barrier.
//
// Layout of W between passes: after a forward row transform the x axis is in
// bit-reversed order, after a forward column transform the y axis too; the
// inverse transforms consume bit-reversed input, so no permutation pass
// exists anywhere.  The measured intensities and transfer functions are
// permuted once at plan creation to match (transposed, bit-reversed) so the
// column pass reads them coalesced.
//
// Numerics follow the reference's complex64 CPU path operation by operation
// (BaseEngine.py:870-1012 projection, ePIE.py:128-160, mPIE.py:196-270,
// multiPIE.py:125-205, e3PIE.py:89-225); the per-position error is summed in
// FP64 and divided by the FP32 frame energy as getErrorMetrics does
// (BaseEngine.py:819-841).

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "pty_fft.h"
#include "pty_hip.h"

using namespace pty;

static_assert(sizeof(pty_instance) == 216, "pty_instance layout (ABI 4) changed: update the bindings");

#define NTHREADS 256
#define NWAVE (NTHREADS / 64)
#define MAX_SLICES 15
#define SPIN_LIMIT (1u << 22)

// control block of an instance (32-bit words)
#define CTRL_TOP 0       // counter barrier: arrivals of complete groups
#define CTRL_FAIL 1      // set when a cluster barrier timed out (every workgroup then leaves)
#define CTRL_PLACE 2     // set when a workgroup of an XCD-local cluster found no slot on its XCD
#define CTRL_BADPOS 3    // set when a position (rewritten by the caller between launches) left the object
#define CTRL_GROUP 32    // counter barrier: 8 group arrival counters, 32 words apart
#define CTRL_RELEASE 288 // counter barrier: published generation
#define CTRL_XPOP 320    // XCD barrier: workgroups of the cluster on each XCD (8 words)
#define CTRL_XCTR 352    // XCD barrier: arrivals on each XCD, 32 words apart, counted in that XCD's L2
#define CTRL_XTOP 608    // XCD barrier: arrivals of complete XCDs
#define CTRL_XREL 640    // XCD barrier: published generation
#define CTRL_WORDS 672
#define SLOT_WORDS 32   // words of per-position maxima in each of the 4 ring slots
#define SLOT_QN_OMAX 2   // qNewton's object maximum (words 0, 1: the first row pass's maxima)

#define N_XCD 8  // XCDs of an MI355X (gfx950), each with its own L2
#define PC_TERMS 19  // position correction: 9 complex correlations and the patch norm

// pass families of the persistent kernel
#define FAM_PIE_FRAUNHOFER 0
#define FAM_PIE_ASP 1
#define FAM_E3PIE 2
#define FAM_QN_FRAUNHOFER 3  // qNewton: its own instantiations keep the PIE kernels' registers
#define FAM_QN_ASP 4
#define FAM_PIE_SASP 5  // PIE with scaled ASP: the ASP passes plus a phase in the row passes
// PIE families with the position correction pass (their own instantiations,
// so its FP64 accumulators stay out of the other kernels' registers); per
// probe size only, never in the batched launch
#define FAM_PC_FRAUNHOFER 6
#define FAM_PC_ASP 7
#define FAM_PC_SASP 8
#define N_FAMILIES 9
__host__ __device__ constexpr int base_family(int fam) {
  return fam == FAM_PC_FRAUNHOFER ? FAM_PIE_FRAUNHOFER : fam == FAM_PC_ASP ? FAM_PIE_ASP : fam == FAM_PC_SASP ? FAM_PIE_SASP : fam;
}

namespace {

thread_local std::string g_err;
int set_err(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int set_err(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return set_err(PTY_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));      \
  } while (0)

}  // namespace

// ---------------------------------------------------------------------------
// device-side instance descriptor
// ---------------------------------------------------------------------------
struct DevInst {
  pty_instance a;
  int G;             // workgroups of this instance
  int nF;            // fields in W
  int nFp;           // nF rounded up to a power of two (item slots per line)
  int fs;            // slice stride of W fields (last-slice fields: f % fs == fs-1)
  int nO, nPr;       // object / probe fields
  float omAlphaO;    // (float)(1 - alphaObject), rounded as NumPy does
  float omAlphaP;
  float2* W;         // (nF, Np, Np)
  float2* ESW;       // e3PIE exit waves (nslice, nF, Np, Np)
  const float* Ik;   // ptychogram in kernel layout
  const float2* HT;  // transfer function(s), bit-reversed layout, / Np^2; one per wavelength
  const float2* Q;   // Fresnel quadratic phase (Np, Np), natural order, or null
  const unsigned char* live;  // ASP: per transfer function, 1 where column p of HT has a nonzero entry; or null
  const short* lgrp;  // ASP: per transfer function, the column groups (of the kernel's width) with a live column
  const int* nlive;   // [0]: most live groups of any transfer function, [1 + t]: those of function t; or null
  const short* cix;   // ASP families without qNewton: per transfer function and column, where the column sits in
                      // the packed exchange rows (see load_row_live), or null when W keeps full rows
  int wst;            // row stride of the packed exchange rows (live groups x column width), when cix is set
  int fpl;           // W fields per transfer function (fields of one wavelength)
  float2* OP;        // qNewton: the object patch before the position's update (nO, Np, Np)
  unsigned* ctrl;    // control block, CTRL_* layout
  unsigned* mx;      // per-position maxima: 4 ring slots x SLOT_WORDS words x G workgroup partials
  unsigned* bar;     // flag barrier: generation reached by each workgroup (G words)
  double* adampart;  // mqNewton: per-workgroup FP64 partials of |gradient|^2, (2, G)
  double* adamv;     // mqNewton: ADAM second moments of object and probe, (2 parities, 2)
  double* pcpart;    // position correction: per-workgroup FP64 partials, (G, PC_TERMS)
  float2* PCW;       // Fourier position correction: both patches' transforms, (2, Np, Np)
  double* pcc;       // Fourier position correction: |xcor| on the (2r + 1)^2 shifts
  unsigned osel[32]; // fields whose object terms the owner of field f's object sums (bit g)
  unsigned psel[32]; // fields whose probe terms the owner of field f's probe sums (bit g)
  int fl;            // the row passes may loop over the fields of a line in one wave: one
                     // object field summing every field, probe field f taking field f alone
  int gsize[8];      // counter barrier: workgroups of this instance in each block group b % 8
  int ngroups;       // counter barrier: non-empty groups
  double* errpart;   // (2, F, G) per-position per-workgroup error partials
};

__device__ float2 g_twiddle[6][256];  // exp(-2 pi i j / N), j < N/2, N = 16..512

// propagators whose detector transform folds fft2c's shifts into a checkerboard
__host__ __device__ static inline bool prop_checkerboard(int prop) { return prop == PTY_FRAUNHOFER || prop == PTY_FRESNEL; }
// propagators of the angular-spectrum family (detector = ifft(fft(.) H))
__host__ __device__ static inline bool prop_asp(int prop) {
  return prop == PTY_ASP || prop == PTY_POLYCHROME_ASP || prop == PTY_SCALED_ASP || prop == PTY_SCALED_POLYCHROME_ASP;
}

__host__ __device__ static int log2_np(int np) {
  switch (np) {
    case 16: return 4;
    case 32: return 5;
    case 64: return 6;
    case 128: return 7;
    case 256: return 8;
    case 512: return 9;
  }
  return -1;
}

// threads per workgroup of the per-size kernel (Cfg<NP>::NT) and item slots
// per round for a workgroup of nt threads: waves x lines per wave
static int nthreads_rt(int np) { return np >= 128 ? 512 : 1024; }
static int slots_rt(int np, int nt) { return (nt / 64) * (np >= 64 ? 1 : 64 / np); }

struct Run {
  const DevInst* D;
  int g, G, wave, t, sub, lane;
  int nF, nFp, nwave;
  int nt, ipr;  // threads per workgroup, (line, field) item slots per round
  unsigned gen;
  const float2* tw;
  double* red;
  float2* xa;  // LDS exchange buffers, one line-field slot per item
  float2* xb;
  float2* fx;  // this wave's LDS scratch of the transposed line transforms
  int* fail;
  unsigned* smx;  // LDS: the workgroup's running maxima and arrival count of publish_max2
  bool loc;  // every workgroup of the cluster runs on one XCD: exchange through its L2
  bool fl;   // field-loop row passes (DevInst::fl, and at least one line per item slot)
  bool xbar;  // XCD barrier (cluster spread over the XCDs; xcd_barrier_setup done)
  unsigned xgen;  // XCD barrier: generation
  unsigned xcd, xpop, nx;  // thread 0: its XCD, the cluster's workgroups there, non-empty XCDs
  const short* lcix;  // LDS copy of DevInst::cix, or null
};

// ---------------------------------------------------------------------------
// reductions and the cluster barrier
// ---------------------------------------------------------------------------
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, xlane(v, m));
  return v;
}
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += xlane_d(v, m);
  return v;
}

// Block-wide FP64 sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(Run& R, double v) {
  v = wave_sum(v);
  __syncthreads();
  if (R.lane == 0) R.red[R.wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < R.nwave; ++w) s += R.red[w];
  }
  return s;
}

// Per-position maxima (of non-negative floats, ordered as their bit
// patterns).  Each workgroup keeps its own partial of every word of a ring
// slot, at slot + w * G + g, which its waves combine with an atomic max; a
// reader takes the maximum over the G partials.  Atomics on G separate
// addresses proceed in parallel where one shared word per maximum would
// serialise every wave of the cluster.
__device__ __forceinline__ void atomic_fmax_slot(const Run& R, unsigned* slot, int w, float v) {
  __hip_atomic_fetch_max(slot + w * R.G + R.g, __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// called by every lane of a wave; the maxima of words w0 and w1 over the G
// partials, returned to all of them.  Every load of both words is issued
// before the first is waited for, so the pair costs one round trip for
// G <= 256 (a lane past the last partial loads the last one again: the
// maximum is the same).
__device__ __forceinline__ void load_slot2(const Run& R, const unsigned* slot, int w0, int w1, float& m0,
                                           float& m1) {
  const unsigned* p0 = slot + w0 * R.G;
  const unsigned* p1 = slot + w1 * R.G;
  const int last = R.G - 1;
  float a = 0.f, b = 0.f;
  for (int base = 0; base < R.G; base += 256) {
    unsigned va[4], vb[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i0 = base + 64 * u + R.lane;
      const int i = i0 < last ? i0 : last;
      va[u] = __hip_atomic_load(p0 + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      vb[u] = __hip_atomic_load(p1 + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a = fmaxf(a, __uint_as_float(va[u]));
      b = fmaxf(b, __uint_as_float(vb[u]));
    }
  }
  // uniform: kept in scalar registers
  m0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(a))));
  m1 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(b))));
}
// one word, in the same way
__device__ __forceinline__ float load_slot(const Run& R, const unsigned* slot, int w) {
  const unsigned* p = slot + w * R.G;
  const int last = R.G - 1;
  float a = 0.f;
  for (int base = 0; base < R.G; base += 256) {
    unsigned v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int i0 = base + 64 * u + R.lane;
      v[u] = __hip_atomic_load(p + (i0 < last ? i0 : last), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) a = fmaxf(a, __uint_as_float(v[u]));
  }
  return __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(a))));
}
// clears this workgroup's partials of a ring slot
__device__ __forceinline__ void clear_slot(const Run& R, unsigned* slot) {
  if (threadIdx.x < SLOT_WORDS) __hip_atomic_store(slot + threadIdx.x * R.G + R.g, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// clears this workgroup's partial of one word of a ring slot
__device__ __forceinline__ void clear_slot_word(const Run& R, unsigned* slot, int w) {
  if (threadIdx.x == 0) __hip_atomic_store(slot + w * R.G + R.g, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The first row pass's maxima (probe word 0, object word 1): the waves of a
// workgroup combine theirs with LDS atomics, and the last wave to finish
// publishes the workgroup's partials with plain stores (device-coherent ones
// across XCDs).  Every workgroup writes its partials at every position, so
// these words need no clearing, and no global atomic read-modify-write is
// issued: its round trip past the L2 is what the barrier's store drain would
// otherwise wait for.  The LDS words are reset by the publishing wave; the
// next position's pass is several workgroup barriers later.
__device__ __forceinline__ void store_slot(const Run& R, unsigned* p, unsigned v) {
  if (R.loc)
    *(__attribute__((address_space(1))) unsigned*)p = v;
  else
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void publish_max2(const Run& R, unsigned* slot, float a, float b) {
  a = wave_max(a);
  b = wave_max(b);
  if (R.lane == 0) {
    unsigned* m = R.smx;
    __hip_atomic_fetch_max(m, __float_as_uint(a), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_max(m + 1, __float_as_uint(b), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const unsigned n = __hip_atomic_fetch_add(m + 2, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (n + 1u == (unsigned)R.nwave) {
      const unsigned pa = __hip_atomic_load(m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      const unsigned pb = __hip_atomic_load(m + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(m, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(m + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      __hip_atomic_store(m + 2, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      store_slot(R, slot + R.g, pa);
      store_slot(R, slot + R.G + R.g, pb);
    }
  }
}

// Cross-workgroup data of the persistent kernel (exit waves, object, probe,
// momenta, maxima, error partials) is only ever read and written with
// coherent accesses: agent-scope relaxed atomics (global_load/store ... sc1)
// in general, or, for a cluster on one XCD, plain stores and sc1 loads that
// the XCD's L2 serves (below); no cache the readers use holds a stale copy.
// A barrier then needs no cache maintenance: every wave drains its
// outstanding stores (s_waitcnt vmcnt(0): the stores have been performed at
// the coherence point), the workgroup barrier orders that before the
// workgroup's arrival, and once every workgroup has arrived the cluster
// proceeds.  This avoids the L2 write-back/invalidate of agent-scope fences,
// whose cost grows with the number of workgroups per XCD.  The spin is
// bounded: on time-out the failure flag is raised and every workgroup of the
// instance leaves.
__device__ __forceinline__ void drain_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */ }

// XCD-local accessors.  When every workgroup of a cluster runs on one XCD,
// cross-workgroup data need only be coherent in that XCD's L2: stores are
// plain (the CU's L1 is write-through, so a drained store is in L2) and
// loads are sc1 (device-scope) buffer loads, which the L1 never serves and
// the L2 does (about 215 cycles on a hit, against 108 for an L1 hit and 560
// for a line from memory).  An sc0 load may hit a line the CU's L1 cached
// before another workgroup rewrote it (Np = 32 clusters of 4 and 8
// workgroups read stale exit-wave lines that way); the sc1 form needs no L1
// invalidate at the barrier, which cost 0.8-1.1 us per barrier.  The buffer
// resource is built from the first active lane's address minus 2^31, so the
// 32-bit offsets of one wave's accesses reach 2 GiB either side of it (a
// wave's accesses of a pass lie within one field).
#define LOC_AUX 16  // cache policy of the XCD-local loads: sc1
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
struct LocalRef {
  __amdgpu_buffer_rsrc_t r;
  uint32_t off;  // this lane's byte offset from the resource base
};
__device__ __forceinline__ LocalRef local_ref(const void* p) {
  const uint64_t a = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
  uint64_t base = ((uint64_t)hi << 32) | lo;
  base = base >= 0x80000000ull ? base - 0x80000000ull : 0ull;
  return {__builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, -1, 0x00020000), (uint32_t)(a - base)};
}
__device__ __forceinline__ unsigned wave_min_u32(unsigned v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    const unsigned o = (unsigned)xlane_i((int)v, m);
    v = o < v ? o : v;
  }
  return v;
}

// partial i of a maxima word (XCD-local or device-coherent load)
__device__ __forceinline__ unsigned load_partial(const Run& R, const unsigned* p, int i) {
  if (R.loc) {
    const LocalRef L = local_ref(p);
    return __builtin_amdgcn_raw_buffer_load_b32(L.r, L.off + (uint32_t)i * 4u, 0, LOC_AUX);
  }
  return __hip_atomic_load(p + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// called by every lane of a wave; the maxima of the G partials of words 0
// and 1 that publish_max2 wrote, returned to all of them.  Every load of both
// words is issued before the first is waited for, so the pair costs one
// round trip for G <= 256 (a lane past the last partial loads the last one
// again: the maximum is the same).
__device__ __forceinline__ void load_max_published2(const Run& R, const unsigned* slot, float& m0, float& m1) {
  const unsigned* p0 = slot;
  const unsigned* p1 = slot + R.G;
  const int last = R.G - 1;
  float a = 0.f, b = 0.f;
  if (R.G <= 64) {
    const int i = R.lane < last ? R.lane : last;
    a = __uint_as_float(load_partial(R, p0, i));
    b = __uint_as_float(load_partial(R, p1, i));
  } else {
    for (int base = 0; base < R.G; base += 256) {
      unsigned va[4], vb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int i0 = base + 64 * u + R.lane;
        const int i = i0 < last ? i0 : last;
        va[u] = load_partial(R, p0, i);
        vb[u] = load_partial(R, p1, i);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a = fmaxf(a, __uint_as_float(va[u]));
        b = fmaxf(b, __uint_as_float(vb[u]));
      }
    }
  }
  m0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(a))));
  m1 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(b))));
}
// load_max_published2 in two halves, for a pass that reads the maxima for a
// later one: max_issue issues the loads (MU partials per lane and word, for
// G <= 64 MU), max_finish reduces them (a larger cluster reads them all
// there)
template <int MU = 4>
struct MaxLoads {
  unsigned a[MU], b[MU];
};
template <int MU>
__device__ __forceinline__ void max_issue(const Run& R, const unsigned* slot, MaxLoads<MU>& m) {
  const int last = R.G - 1;
#pragma unroll
  for (int u = 0; u < MU; ++u) {
    const int i0 = 64 * u + R.lane;
    const int i = i0 < last ? i0 : last;
    m.a[u] = R.G <= 64 * MU ? load_partial(R, slot, i) : 0u;
    m.b[u] = R.G <= 64 * MU ? load_partial(R, slot + R.G, i) : 0u;
  }
}
template <int MU>
__device__ __forceinline__ void max_finish(const Run& R, const unsigned* slot, const MaxLoads<MU>& m, float& m0,
                                           float& m1) {
  if (R.G > 64 * MU) {
    load_max_published2(R, slot, m0, m1);
    return;
  }
  float a = 0.f, b = 0.f;
#pragma unroll
  for (int u = 0; u < MU; ++u) {
    a = fmaxf(a, __uint_as_float(m.a[u]));
    b = fmaxf(b, __uint_as_float(m.b[u]));
  }
  m0 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(a))));
  m1 = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(wave_max(b))));
}

// Arrival is two-level: a workgroup increments the counter of its XCD group
// (blocks of one launch are dispatched round-robin, block b on XCD b % 8;
// the grouping only spreads the atomics over separate cache lines, nothing
// depends on the placement), the last arriver of a group increments the top
// counter, and the last arriver overall publishes the generation on a
// release line that the waiting workgroups poll.  Arrivals and polls thus
// never contend for one address.
__device__ __forceinline__ void counter_sync(Run& R) {
  if (threadIdx.x == 0) {
    const DevInst& D = *R.D;
    unsigned* ctrl = D.ctrl;
    const int grp = blockIdx.x & 7;
    unsigned* gctr = ctrl + CTRL_GROUP + 32 * grp;
    unsigned* top = ctrl + CTRL_TOP;
    unsigned* rel = ctrl + CTRL_RELEASE;
    const unsigned gen = R.gen;
    const unsigned old = __hip_atomic_fetch_add(gctr, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old + 1u == gen * (unsigned)D.gsize[grp]) {
      const unsigned t = __hip_atomic_fetch_add(top, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (t + 1u == gen * (unsigned)D.ngroups) __hip_atomic_store(rel, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned spins = 0;
    int fail = 0;
    while (__hip_atomic_load(rel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen) {
      ++spins;
      if ((spins & 255u) == 0u &&
          (spins > SPIN_LIMIT || __hip_atomic_load(ctrl + CTRL_FAIL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(ctrl + CTRL_FAIL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *R.fail = fail;
  }
}

// XCD barrier, for a cluster spread over the XCDs: a workgroup arrives on
// the counter of the XCD it runs on with an atomic performed in that XCD's L2
// (workgroup scope: no round trip past the L2), the last arriver of an XCD
// increments the device-scope top counter, and the last XCD to complete
// publishes the generation on a release line that every workgroup polls.
// Against the counter barrier above, one of the two serial device-scope
// atomic round trips on the arrival path becomes an L2 atomic.  The
// exchange data themselves are written and read with device-scope accesses
// and drained before the arrival, so they are visible across XCDs once the
// release is seen.  Each workgroup reads its XCD from the XCC_ID register;
// the counts per XCD come from xcd_barrier_setup, so nothing depends on how
// the dispatcher spreads the workgroups.
__device__ __forceinline__ unsigned xcc_id() {
  return (unsigned)__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7u;
}
__device__ __forceinline__ void xcd_sync(Run& R) {
  if (threadIdx.x == 0) {
    unsigned* ctrl = R.D->ctrl;
    const unsigned gen = R.xgen;
    const unsigned old =
        __hip_atomic_fetch_add(ctrl + CTRL_XCTR + 32 * R.xcd, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old + 1u == gen * R.xpop) {
      const unsigned t = __hip_atomic_fetch_add(ctrl + CTRL_XTOP, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (t + 1u == gen * R.nx) __hip_atomic_store(ctrl + CTRL_XREL, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    unsigned spins = 0;
    int fail = 0;
    while (__hip_atomic_load(ctrl + CTRL_XREL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < gen) {
      ++spins;
      if ((spins & 255u) == 0u &&
          (spins > SPIN_LIMIT || __hip_atomic_load(ctrl + CTRL_FAIL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        __hip_atomic_store(ctrl + CTRL_FAIL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    *R.fail = fail;
  }
}

// Flag barrier: workgroup g publishes the generation it has reached in its
// own word D.bar[g] (a store, no atomic); wave 0 of every workgroup polls
// the G words until all of them reach the generation.  Used by clusters on
// one XCD: stores are plain and the polls device-scope atomic loads (sc1),
// which the XCD's L2 serves without the L1 and the compiler keeps inside the
// loop.  (System-scope polls, sc0 sc1, took 1.7 us for one round of loads
// where a word just written by other CUs was read.)  Measured against the
// counter barrier above it is the faster one within an XCD and the slower
// one across XCDs, where every poll of G words goes past the L2s.
__device__ __forceinline__ void flag_sync(Run& R) {
  if (R.wave == 0) {
    const DevInst& D = *R.D;
    unsigned* bar = D.bar;
    unsigned* ctrl = D.ctrl;
    const unsigned gen = R.gen;
    if (R.lane == 0) {
      *(__attribute__((address_space(1))) unsigned*)(bar + R.g) = gen;
    }
    unsigned spins = 0;
    int fail = 0;
    for (;;) {
      unsigned mn = 0xffffffffu;
      for (int g = R.lane; g < R.G; g += 64) {
        const unsigned v = __hip_atomic_load(bar + g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mn = v < mn ? v : mn;
      }
      if (wave_min_u32(mn) >= gen) break;
      ++spins;
      if ((spins & 255u) == 0u &&
          (spins > SPIN_LIMIT || __hip_atomic_load(ctrl + CTRL_FAIL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
        if (R.lane == 0) __hip_atomic_store(ctrl + CTRL_FAIL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fail = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (R.lane == 0) *R.fail = fail;
  }
}

// coherent (agent-scope) accessors for cross-workgroup data
__device__ __forceinline__ float2 cload(const float2* p) {
  const uint64_t u = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __builtin_bit_cast(float2, u);
}
__device__ __forceinline__ void cstore(float2* p, float2 v) {
  __hip_atomic_store((uint64_t*)p, __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double cloadd(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cstored(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void cstoreu(unsigned* p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The cluster barrier: the flag form for a cluster on one XCD, the
// counter form otherwise.
__device__ __forceinline__ bool cluster_sync(Run& R) {
  drain_stores();
  __syncthreads();
  if (R.G == 1) return true;
  R.gen += 1u;
  if (R.loc) {
    flag_sync(R);
  } else if (R.xbar) {
    R.xgen += 1u;
    xcd_sync(R);
  } else {
    counter_sync(R);
  }
  __syncthreads();
  return *R.fail == 0;
}

// run-state accessors: XCD-local or device-coherent (local_ref above)
__device__ __forceinline__ float2 lload(const LocalRef& L, uint32_t byte_off) {
  const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(L.r, L.off + byte_off, 0, LOC_AUX);
  return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
}
__device__ __forceinline__ float2 cload(const Run& R, const float2* p) {
  if (R.loc) return lload(local_ref(p), 0u);
  return cload(p);
}
__device__ __forceinline__ void cstore(const Run& R, float2* p, float2 v) {
  if (R.loc)
    *(__attribute__((address_space(1))) uint64_t*)p = __builtin_bit_cast(uint64_t, v);
  else
    cstore(p, v);
}
__device__ __forceinline__ double cloadd(const Run& R, const double* p) {
  if (R.loc) {
    const LocalRef L = local_ref(p);
    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(L.r, L.off, 0, LOC_AUX);
    return __builtin_bit_cast(double, w);
  }
  return cloadd(p);
}
__device__ __forceinline__ void cstored(const Run& R, double* p, double v) {
  if (R.loc)
    *(__attribute__((address_space(1))) double*)p = v;
  else
    cstored(p, v);
}

// ---------------------------------------------------------------------------
// line helpers
// ---------------------------------------------------------------------------
// Global-address-space views: pointers read from the descriptor are generic
// pointers to the compiler; casting them back lets it emit global (not flat)
// memory instructions.
#define GAS __attribute__((address_space(1)))

template <int NP>
__device__ __forceinline__ void load_row(const Run& R, float2 (&v)[Line<NP>::L], const float2* row, int t) {
  if (R.loc) {
    const LocalRef L = local_ref(row + t);
#pragma unroll
    for (int l = 0; l < Line<NP>::L; ++l) v[l] = lload(L, (uint32_t)(l * Line<NP>::LPF * 8));
    return;
  }
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) v[l] = cload(row + l * Line<NP>::LPF + t);
}
template <int NP>
__device__ __forceinline__ void store_row(const Run& R, float2* row, const float2 (&v)[Line<NP>::L], int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) cstore(R, row + l * Line<NP>::LPF + t, v[l]);
}
// device-coherent forms for the standalone propagation kernels
template <int NP>
__device__ __forceinline__ void load_row(float2 (&v)[Line<NP>::L], const float2* row, int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) v[l] = cload(row + l * Line<NP>::LPF + t);
}
template <int NP>
__device__ __forceinline__ void store_row(float2* row, const float2 (&v)[Line<NP>::L], int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) cstore(row + l * Line<NP>::LPF + t, v[l]);
}
template <int NP>
__device__ __forceinline__ void load_col(float2 (&v)[Line<NP>::L], const float2* f, int p, int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) v[l] = cload(f + (size_t)(l * Line<NP>::LPF + t) * NP + p);
}
template <int NP>
__device__ __forceinline__ void store_col(float2* f, const float2 (&v)[Line<NP>::L], int p, int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) cstore(f + (size_t)(l * Line<NP>::LPF + t) * NP + p, v[l]);
}
// Frequency-side rows: the forward row transform leaves, and the inverse
// takes, register l of lane t at slot rslot(l, t) of the row (the transposed
// form's last layout; fft_fwd / fft_inv with Z); storage keeps slot order.
template <int NP>
__device__ __forceinline__ void load_row_f(const Run& R, float2 (&v)[Line<NP>::L], const float2* row, int t) {
  if (R.loc) {
    const int s0 = rslot<NP>(0, t);
    const LocalRef L = local_ref(row + s0);
#pragma unroll
    for (int l = 0; l < Line<NP>::L; ++l) v[l] = lload(L, (uint32_t)((rslot<NP>(l, t) - s0) * 8));
    return;
  }
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) v[l] = cload(row + rslot<NP>(l, t));
}
template <int NP>
__device__ __forceinline__ void store_row_f(const Run& R, float2* row, const float2 (&v)[Line<NP>::L], int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) cstore(R, row + rslot<NP>(l, t), v[l]);
}
// read-only inputs (measured frames, transfer function): ordinary cached loads
__device__ __forceinline__ float ldg(const float* p) { return *(const GAS float*)p; }
__device__ __forceinline__ float2 ldg2(const float2* p) {
  return __builtin_bit_cast(float2, *(const GAS unsigned long long*)p);
}

// Row elements in columns that the column pass skipped (outside the band
// limit of field f's transfer function) are zero.  live_bits reads the
// column flags of this lane's frequency-side elements (bit l for slot
// rslot(l, t)); it is issued before the row loads so that its latency hides
// under theirs.
template <int NP>
__device__ __forceinline__ unsigned live_bits(const DevInst& D, int f, int t) {
  if (!D.live) return ~0u;
  const unsigned char* lv = D.live + (size_t)(f / D.fpl) * NP;
  unsigned b = 0u;
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) b |= (lv[rslot<NP>(l, t)] ? 1u : 0u) << l;
  return b;
}
template <int NP>
__device__ __forceinline__ void zero_dead_columns(unsigned bits, float2 (&v)[Line<NP>::L]) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l)
    if (!((bits >> l) & 1u)) v[l] = make_float2(0.f, 0.f);
}

// Packed exchange rows.  Between a row pass and a column pass of an
// angular-spectrum propagation only the live column groups (those the
// column passes transform, DevInst::lgrp) carry data; with DevInst::cix set
// W holds just those, packed: row y of field f starts at W + f * NP * ws +
// y * ws (ws = DevInst::wst), and the m-th live group of field f's transfer
// function sits at columns [m * CW, (m + 1) * CW).  cix[p], per column p of
// the row layout: the packed column when column p is live; -1 - the packed
// column when p is dead inside a live group (stored, for the column pass,
// but read as zero by the row passes, as zero_dead_columns does); PACK_DEAD
// when p's group is dead (neither stored nor read).  A row pass thus moves
// the live groups only (33 of 256 columns at the benchmark's band limit)
// instead of full rows; the values the passes compute are unchanged.
#define PACK_DEAD (-32768)
#define PACK_LDS 2048  // column codes a per-size workgroup keeps in LDS (transfer functions x Np)
__device__ __forceinline__ const short* packed_ix(const Run& R, const DevInst& D, int f) {
  if (R.lcix) return R.lcix + (size_t)(f / D.fpl) * D.a.Np;
  return D.cix ? D.cix + (size_t)(f / D.fpl) * D.a.Np : nullptr;
}
// row stride of the frequency-domain exchange rows of W
template <int NP>
__device__ __forceinline__ int xrow_stride(const DevInst& D) { return D.cix ? D.wst : NP; }
// loads a frequency-domain row: full row with dead columns zeroed when ix is
// null (lb from live_bits), else the live columns of the packed row
template <int NP>
__device__ __forceinline__ void load_row_live(const Run& R, float2 (&v)[Line<NP>::L], const float2* row,
                                              const short* ix, unsigned lb, int t) {
  if (!ix) {
    load_row_f<NP>(R, v, row, t);
    zero_dead_columns<NP>(lb, v);
    return;
  }
  short c[Line<NP>::L];
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) c[l] = ix[rslot<NP>(l, t)];
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) v[l] = c[l] >= 0 ? cload(R, row + c[l]) : make_float2(0.f, 0.f);
}
template <int NP>
__device__ __forceinline__ void store_row_live(const Run& R, float2* row, const float2 (&v)[Line<NP>::L],
                                               const short* ix, int t) {
  if (!ix) {
    store_row_f<NP>(R, row, v, t);
    return;
  }
  short c[Line<NP>::L];
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) c[l] = ix[rslot<NP>(l, t)];
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l)
    if (c[l] != PACK_DEAD) cstore(R, row + (c[l] >= 0 ? c[l] : -1 - c[l]), v[l]);
}

// ---------------------------------------------------------------------------
// work decomposition
// ---------------------------------------------------------------------------
// A workgroup owns lines [g * Np/G, (g+1) * Np/G) of every pass.  Its waves
// take (line, field) items: item slot s of a round is line s / nFp, field
// s % nFp (nFp = nF rounded up to a power of two; padding slots compute on a
// valid field and store nothing).  One wave transforms one field of one line
// per round, so the four probe modes of a position run on four waves at once.
// Sums over fields (the modelled intensity, the object and probe increments,
// the maxima) go through LDS and are added in increasing field order, the C
// order of NumPy's axis sums, so results are deterministic.
template <int NP, int CWR = 0>
struct Col {
#ifndef PTY_COLUMN_WIDTH
  // measured on Np = 256 clusters of 128 workgroups: CW = 1 41.2 ms (ASP
  // with band-limited columns skipped) and 32.1 ms (Fraunhofer) per
  // 1000-position iteration, CW = 2 46.3 and 32.3, CW = 4 51.6 (ASP)
  static constexpr int CWD = NP >= 512 ? 2 : 1;
#else
  static constexpr int CWD = PTY_COLUMN_WIDTH;
#endif
  static constexpr int CW0 = CWR > 0 ? CWR : CWD;
  static constexpr int CW = CW0 * NP >= 64 ? CW0 : 64 / NP;  // a column spans at most NP lanes
  static constexpr int LPC = 64 / CW;                         // lanes per column
  static constexpr int CL = NP / LPC;                         // values per lane
  static constexpr int NG = NP / CW;                          // column groups
};

// Col<NP, CWR>::CW of a run-time probe size
template <int CWR>
constexpr int col_width(int np) {
  return np == 16 ? Col<16, CWR>::CW : np == 32 ? Col<32, CWR>::CW : np == 64 ? Col<64, CWR>::CW
       : np == 128 ? Col<128, CWR>::CW : np == 256 ? Col<256, CWR>::CW : Col<512, CWR>::CW;
}

template <int NP>
struct Cfg {
  static constexpr int NT = NP >= 128 ? 512 : 1024;  // threads per workgroup
  static constexpr int NW = NT / 64;
  static constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF, SUB = Line<NP>::SUB;
  static constexpr int IPR = NW * SUB;  // item slots per round
  // LDS exchange: row passes use two (IPR, NP) complex buffers, column
  // passes one (NW, values per lane, 64) float buffer
  static constexpr int XROW = 2 * IPR * NP;
  static constexpr int XCOL = NW * NP * Col<NP>::CW / 2;
  static constexpr int XSTG = NW * Col<NP>::CW * (2 * NP + 32) * (NP <= 256);  // staged column rounds
  static constexpr int XRC = XROW > XCOL ? XROW : XCOL;
  static constexpr int XBUF = XRC > XSTG ? XRC : XSTG;  // float2 elements
};

#define ITEMS_BEGIN(R, NP, line, f, act, islot, lslot, frw)                     \
  {                                                                            \
    const int per_ = NP / (R).G;                                               \
    const int line0_ = (R).g * per_;                                           \
    const int lpr_ = (R).ipr / (R).nFp;                                        \
    const int islot = (R).wave * Line<NP>::SUB + (R).sub;                      \
    const int lil_ = islot / (R).nFp;                                          \
    const int frw = islot - lil_ * (R).nFp;                                    \
    const int lslot = lil_ * (R).nFp;                                          \
    const int f = frw < (R).nF ? frw : (R).nF - 1;                             \
    for (int base_ = 0; base_ < per_; base_ += lpr_) {                         \
      const int li_ = base_ + lil_;                                            \
      const bool lok_ = li_ < per_;                                            \
      const bool act = lok_ && frw < (R).nF;                                   \
      const int line = line0_ + (lok_ ? li_ : per_ - 1);
#define ITEMS_END \
  }               \
  }
#define ITEM_LEAD(act, frw) ((act) && (frw) == 0)  // the slot that speaks for its line
// End of a round whose LDS slots the next round rewrites.  After the last
// round the barrier that follows the pass (cluster barrier or block_sum)
// orders the reuse, so the workgroup barrier is skipped there.
#define ITEM_ROUND_SYNC() \
  if (base_ + lpr_ < per_) __syncthreads()

// ---------------------------------------------------------------------------
// field bookkeeping
// ---------------------------------------------------------------------------
// Exit-wave field f of the PIE engines enumerates (lambda, osm, psm, slice) in
// C order; its object field is (lambda, osm, slice) and its probe field
// (lambda, psm, slice), matching the 6D layouts.
struct FieldIdx {
  int il, io, ip, is, fo, fp;
};
__host__ __device__ __forceinline__ FieldIdx field_idx(int f, int no, int npm, int ns) {
  FieldIdx x;
  x.is = f % ns;
  x.ip = (f / ns) % npm;
  x.io = (f / (ns * npm)) % no;
  x.il = f / (ns * npm * no);
  x.fo = (x.il * no + x.io) * ns + x.is;
  x.fp = (x.il * npm + x.ip) * ns + x.is;
  return x;
}

template <int NP>
__device__ __forceinline__ void lds_put(float2* buf, int islot, const float2 (&v)[Line<NP>::L], int t) {
#pragma unroll
  for (int l = 0; l < Line<NP>::L; ++l) buf[islot * NP + l * Line<NP>::LPF + t] = v[l];
}

// ---------------------------------------------------------------------------
// ePIE / mPIE / multiPIE position passes
// ---------------------------------------------------------------------------
// First row pass: exit wave psi_f = O_f P_f of the gathered object row and the
// probe row, (checkerboard for the centred Fraunhofer transform), forward row
// transform into W; per pixel sums of |P|^2 over the probe fields and of |O|^2
// over the object fields give the maxima the updates divide by.
template <int NP, bool QN, bool ASPF, bool QPH, bool PK, bool KEEP = false>
__device__ __forceinline__ void pie_rows_first(Run& R, int k, unsigned* slot, unsigned* zslot,
                                               float2 (&ko)[Line<NP>::L], float2 (&kp)[Line<NP>::L]) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int no = D.a.nosm, npm = D.a.npsm, ns = D.a.nslice;
  const int No = D.a.No, nF = D.nF;
  const size_t No2 = (size_t)No * No, Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  const float2* __restrict__ O = (const float2*)D.a.object;
  const float2* __restrict__ P = (const float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  const bool cb = prop_checkerboard(D.a.propagator);
  const float2* __restrict__ Qp = D.Q;
  constexpr bool qn = QN;
  if constexpr (QN) clear_slot_word(R, zslot, SLOT_QN_OMAX);
  float pmax = 0.f, omax = 0.f;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    const FieldIdx x = field_idx(f, no, npm, ns);
    float2 of[L], pf[L], q[L];
    load_row<NP>(R, of, O + x.fo * No2 + (size_t)(row0 + r) * No + col0, R.t);
    load_row<NP>(R, pf, P + x.fp * Np2 + (size_t)r * NP, R.t);
    if constexpr (KEEP) {
      // the last row pass of this position reuses this item's patch and
      // probe rows (no pass in between writes them)
#pragma unroll
      for (int l = 0; l < L; ++l) {
        ko[l] = of[l];
        kp[l] = pf[l];
      }
    }
#pragma unroll
    for (int l = 0; l < L; ++l) {
      // qNewton.py:114: Pmax = max(sum |P|); the PIE engines take sum |P|^2
      const float pv = qn ? sqrtf(cabs2(pf[l])) : cabs2(pf[l]);
      q[l] = make_float2(act && x.io == 0 ? pv : 0.f, act && x.ip == 0 ? cabs2(of[l]) : 0.f);
    }
    // one field: its terms are the field sums, no LDS exchange
    const bool one = nF == 1;
    if (!one) lds_put<NP>(R.xa, islot, q, R.t);
    float2 v[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float2 e = cmul(of[l], pf[l]);
      // Fresnel: fft2c(esw * Q) (Operators.py:110); scaled ASP: c Q1 esw
      // (Operators.py:345); no phase in the other ASP-family kernels
      if constexpr (!ASPF || QPH) {
        // one phase per wavelength (scaledPolychromeASP; fpl = nF otherwise)
        if (Qp) e = cmul(e, ldg2(Qp + (size_t)(f / D.fpl) * Np2 + (size_t)r * NP + l * LPF + R.t));
      }
      if (cb && ((r + l * LPF + R.t) & 1)) e = make_float2(-e.x, -e.y);
      v[l] = e;
    }
    fft_fwd<NP, true>(v, R.tw, R.t, R.fx);
    if constexpr (ASPF) {
      const int ws = PK ? xrow_stride<NP>(D) : NP;
      if (act) store_row_live<NP>(R, W + (size_t)f * NP * ws + (size_t)r * ws, v, (PK ? packed_ix(R, D, f) : nullptr), R.t);
    } else {
      if (act) store_row_f<NP>(R, W + f * Np2 + (size_t)r * NP, v, R.t);
    }
    if (one) {
      if (act) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
          pmax = fmaxf(pmax, 0.f + q[l].x);
          omax = fmaxf(omax, 0.f + q[l].y);
        }
      }
      continue;
    }
    __syncthreads();
    if (ITEM_LEAD(act, frw)) {
      // fields outer, elements inner: the L reads of one field are in
      // flight together (sums still in increasing field order)
      float sp[L], so[L];
#pragma unroll
      for (int l = 0; l < L; ++l) sp[l] = so[l] = 0.f;
      for (int g = 0; g < nF; ++g) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
          const float2 c = R.xa[(lslot + g) * NP + l * LPF + R.t];
          sp[l] += c.x;
          so[l] += c.y;
        }
      }
#pragma unroll
      for (int l = 0; l < L; ++l) {
        pmax = fmaxf(pmax, sp[l]);
        omax = fmaxf(omax, so[l]);
      }
    }
    ITEM_ROUND_SYNC();
  ITEMS_END
  publish_max2(R, slot, pmax, omax);
}

// Column passes: a wave owns a group of CW adjacent columns of one field
// (lane = rho * CW + c, rows n = l * LPC + rho with LPC = 64 / CW lanes per
// column), so every load and store instruction moves contiguous row
// segments of CW * 8 bytes; CW = 1 is a column on 64 adjacent lanes.
// Workgroup g owns column groups [g * cpg, (g + 1) * cpg) of the Np / CW
// groups (workgroups past the last group idle in column passes); its waves
// take (group, field) items, groups per round = waves / nFp.
#define CITEMS_BEGIN(R, NP, p, f, act, islot, lslot, frw)                         \
  {                                                                             \
    const int ng_ = Col<NP, CWR>::NG;                                                \
    const int cpg_ = (ng_ + (R).G - 1) / (R).G;                                 \
    const int cg0_ = (R).g * cpg_;                                              \
    const int cgn_ = cg0_ < ng_ ? (ng_ - cg0_ < cpg_ ? ng_ - cg0_ : cpg_) : 0; \
    const int gpr_ = (R).nwave / (R).nFp;                                       \
    const int islot = (R).wave;                                                 \
    const int lg_ = islot / (R).nFp;                                            \
    const int frw = islot - lg_ * (R).nFp;                                      \
    const int lslot = lg_ * (R).nFp;                                            \
    const int f = frw < (R).nF ? frw : (R).nF - 1;                              \
    for (int base_ = 0; base_ < cgn_; base_ += gpr_) {                          \
      const int gi_ = base_ + lg_;                                              \
      const bool gok_ = gi_ < cgn_;                                             \
      const bool act = gok_ && frw < (R).nF;                                    \
      const int p = (cg0_ + (gok_ ? gi_ : cgn_ - 1)) * Col<NP, CWR>::CW + (R).lane % Col<NP, CWR>::CW;
#define CITEMS_END \
  }                \
  }
#define CITEM_ROUND_SYNC() \
  if (base_ + gpr_ < cgn_) __syncthreads()

template <int NP, int CWR>
__device__ __forceinline__ int col_rho(const Run& R) { return R.lane / Col<NP, CWR>::CW; }

// registers of one lane in a column pass
template <int NP, int CWR>
using ColVals = float2[Col<NP, CWR>::CL];

template <int NP, int CWR>
__device__ __forceinline__ void load_colg(const Run& R, ColVals<NP, CWR>& v, const float2* f, int p, int rho,
                                          int rs = NP) {
  if (R.loc) {
    const LocalRef L = local_ref(f + (size_t)rho * rs + p);
#pragma unroll
    for (int l = 0; l < Col<NP, CWR>::CL; ++l) v[l] = lload(L, (uint32_t)(l * Col<NP, CWR>::LPC * rs * 8));
    return;
  }
#pragma unroll
  for (int l = 0; l < Col<NP, CWR>::CL; ++l) v[l] = cload(f + (size_t)(l * Col<NP, CWR>::LPC + rho) * rs + p);
}
template <int NP, int CWR>
__device__ __forceinline__ void store_colg(const Run& R, float2* f, const ColVals<NP, CWR>& v, int p, int rho,
                                           int rs = NP) {
#pragma unroll
  for (int l = 0; l < Col<NP, CWR>::CL; ++l) cstore(R, f + (size_t)(l * Col<NP, CWR>::LPC + rho) * rs + p, v[l]);
}
template <int NP, int CWR, bool Z = false>
__device__ __forceinline__ void fft_fwd_colg(const Run& R, ColVals<NP, CWR>& v, int rho) {
  fft_fwd_g<NP, Col<NP, CWR>::LPC, Col<NP, CWR>::CW, Z>(v, R.tw, rho, R.fx, R.lane % Col<NP, CWR>::CW);
}
template <int NP, int CWR, bool Z = false>
__device__ __forceinline__ void fft_inv_colg(const Run& R, ColVals<NP, CWR>& v, int rho) {
  fft_inv_g<NP, Col<NP, CWR>::LPC, Col<NP, CWR>::CW, Z>(v, R.tw, rho, R.fx, R.lane % Col<NP, CWR>::CW);
}
// row (of the column line) that register l of lane rho holds on the
// frequency side of fft_fwd_colg / fft_inv_colg with Z
template <int NP, int CWR>
__device__ __forceinline__ int cslot(int l, int rho) {
  return zslot<NP, Col<NP, CWR>::LPC, Col<NP, CWR>::CW>(l, rho);
}

// LDS scratch per wave of the line transforms of probe size NP (rows and
// column groups of the kernel's column width)
template <int NP, int CWR>
constexpr int wave_scratch() { return fft_scratch<NP, Col<NP, CWR>::CW>(); }

// Fraunhofer detector column pass: forward column transform, unitary scale,
// intensity constraint with the measured frame, inverse column transform.
// Fields f with f % fs == fs - 1 form the modelled intensity (the last slice,
// BaseEngine.py:889-891).  Writes this workgroup's FP64 error partial.
template <int NP, int CWR>
__device__ __forceinline__ void detector_cols_fraunhofer_direct(Run& R, int k, int j, double* ep, int fs, float* de) {
  constexpr int CL = Col<NP, CWR>::CL, LPC = Col<NP, CWR>::LPC;
  const DevInst& D = *R.D;
  const size_t Np2 = (size_t)NP * NP;
  float2* __restrict__ W = D.W;
  const float* __restrict__ Ik = D.Ik + (size_t)k * Np2;
  const float invN = 1.0f / (float)NP;
  const int nF = D.nF;
  const int rho = col_rho<NP, CWR>(R);
  float* xs = (float*)R.xa;
  double err = 0.0;
  CITEMS_BEGIN(R, NP, p, f, act, islot, lslot, frw)
    float2 v[CL];
    load_colg<NP, CWR>(R, v, W + f * Np2, p, rho);
    fft_fwd_colg<NP, CWR, true>(R, v, rho);
#pragma unroll
    for (int l = 0; l < CL; ++l) {
      v[l] = cscale(v[l], invN);
      xs[(islot * CL + l) * 64 + R.lane] = act && (f % fs == fs - 1) ? cabs2(v[l]) : 0.f;
    }
    __syncthreads();
    float iest[CL];
#pragma unroll
    for (int l = 0; l < CL; ++l) iest[l] = 0.f;
    for (int g = 0; g < nF; ++g) {
#pragma unroll
      for (int l = 0; l < CL; ++l) iest[l] += xs[((lslot + g) * CL + l) * 64 + R.lane];
    }
#pragma unroll
    for (int l = 0; l < CL; ++l) {
      const int n = cslot<NP, CWR>(l, rho);
      const float im = ldg(Ik + (size_t)n * NP + p);
      if (ITEM_LEAD(act, frw)) {
        err += (double)fabsf(im - iest[l]);
        // element (n, p) is detector pixel (brev(n), brev(p)) (the frames'
        // plan-time permutation)
        if (de) de[(size_t)brev(n, Line<NP>::LOG2) * NP + brev(p, Line<NP>::LOG2)] = fabsf(im - iest[l]);
      }
      v[l] = cscale(v[l], sqrtf(im / (iest[l] + 1e-10f)));
    }
    fft_inv_colg<NP, CWR, true>(R, v, rho);
    if (act) store_colg<NP, CWR>(R, W + f * Np2, v, p, rho);
    CITEM_ROUND_SYNC();
  CITEMS_END
  const double s = block_sum(R, err);
  if (threadIdx.x == 0) cstored(R, ep + (size_t)j * R.G + R.g, s);
}

// The same pass with each round's block staged through LDS.  A round covers
// C = (groups per round) * CW adjacent columns of every field; the workgroup
// loads that block of W (fields x NP rows x C columns) and the C columns of
// the frame with row-contiguous accesses (a load instruction of the direct
// form touches 64 rows, one segment each, here 64 / C rows of C * 8 bytes),
// each wave takes its column group from the tile, and the block goes back
// with row-contiguous stores.  The transforms, sums and constraint are the
// direct form's, so results are bit-identical.  Tiles are column-major with
// a column stride of NP + 16 / C float2 (NP + 32 / C floats for the frame),
// which spreads the rows of one store instruction over distinct LDS banks
// while a wave's column read stays contiguous.  LDS: tiles of at most
// NW * CW * (NP + 16) float2 and NW * CW * (NP + 32) floats plus the
// NW * CL * 64 floats of the field sums: NW * CW * (2 NP + 32) float2 in all
// (XSTG of the kernel configurations), too much for NP = 512.
template <int NP>
constexpr bool col_staged() { return NP <= 256; }
template <int NP, int CWR, int MX = 0>
__device__ __forceinline__ void detector_cols_fraunhofer_staged(Run& R, int k, int j, double* ep, int fs, float* de,
                                                                const unsigned* slot) {
  constexpr int CW = Col<NP, CWR>::CW, CL = Col<NP, CWR>::CL, LPC = Col<NP, CWR>::LPC;
  const DevInst& D = *R.D;
  const size_t Np2 = (size_t)NP * NP;
  float2* __restrict__ W = D.W;
  const float* __restrict__ Ik = D.Ik + (size_t)k * Np2;
  const float invN = 1.0f / (float)NP;
  // the tile geometry is recomputed in every pass (opaque inputs): hoisted
  // out of the position loop it would hold registers through every other pass
  int nF = D.nF, nFp = R.nFp;
  __asm__ volatile("" : "+s"(nF), "+s"(nFp));
  const int rho = col_rho<NP, CWR>(R);
  const int cl = R.lane % CW;
  const int C = (R.nwave / nFp) * CW;  // columns per round, a power of two
  const int lgC = __builtin_ctz((unsigned)C);
  const int lgNC = lgC + Line<NP>::LOG2;
  const int cs = NP + (C == 1 ? 0 : C <= 16 ? 16 / C : 1);
  const int csi = NP + (C == 1 ? 0 : C <= 32 ? 32 / C : 1);
  const int TS = C * cs;
  float2* tile = R.xa;
  float* ikt = (float*)(R.xa + nFp * TS);
  float* xs = ikt + C * csi;
  // MX: wave 0 reads the position's maxima for the update pass here (as in
  // asp_rows_constraint)
  MaxLoads<MX ? MX : 1> mxl;
  if (MX && R.wave == 0) max_issue(R, slot, mxl);
  double err = 0.0;
  CITEMS_BEGIN(R, NP, p, f, act, islot, lslot, frw)
    const int pc0 = (cg0_ + base_) * CW;                               // first column of the round
    const int Cv = (cgn_ - base_ < gpr_ ? cgn_ - base_ : gpr_) * CW;  // its columns
    {
      // every load of the round (its W block and its frame columns, at most
      // CL elements each per thread) is issued before the first LDS write
      // waits for one
      float2 w[CL];
      float iv[CL];
#pragma unroll
      for (int i = 0; i < CL; ++i) {
        const int e = i * R.nt + (int)threadIdx.x;
        const int s = e >> lgNC, n = (e >> lgC) & (NP - 1), c = e & (C - 1);
        const int sc = s < nF ? s : nF - 1, cc = c < Cv ? c : Cv - 1;
        w[i] = cload(R, W + sc * Np2 + (size_t)n * NP + pc0 + cc);
      }
#pragma unroll
      for (int i = 0; i < CL; ++i) {
        const int e = i * R.nt + (int)threadIdx.x;
        const int n = e >> lgC, c = e & (C - 1);
        iv[i] = 0.f;
        if (e < (NP << lgC)) iv[i] = ldg(Ik + (size_t)n * NP + pc0 + (c < Cv ? c : Cv - 1));
      }
#pragma unroll
      for (int i = 0; i < CL; ++i) {
        const int e = i * R.nt + (int)threadIdx.x;
        const int s = e >> lgNC, n = (e >> lgC) & (NP - 1), c = e & (C - 1);
        if (s < nF && c < Cv) tile[s * TS + c * cs + n] = w[i];
      }
#pragma unroll
      for (int i = 0; i < CL; ++i) {
        const int e = i * R.nt + (int)threadIdx.x;
        const int n = e >> lgC, c = e & (C - 1);
        if (e < (NP << lgC)) ikt[c * csi + n] = iv[i];
      }
    }
    __syncthreads();
    // this lane's tile column; a wave past the round's last group repeats
    // that group (and stores nothing), as in the direct form
    const int q = (lg_ * CW < Cv ? lg_ * CW : Cv - CW) + cl;
    float2 v[CL];
#pragma unroll
    for (int l = 0; l < CL; ++l) v[l] = tile[f * TS + q * cs + l * LPC + rho];
    fft_fwd_colg<NP, CWR, true>(R, v, rho);
    float iest[CL];
#pragma unroll
    for (int l = 0; l < CL; ++l) v[l] = cscale(v[l], invN);
    if (nF == 1) {
      // one field: the modelled intensity is its own (0 + x = x)
#pragma unroll
      for (int l = 0; l < CL; ++l) iest[l] = cabs2(v[l]);
    } else {
#pragma unroll
      for (int l = 0; l < CL; ++l) xs[(islot * CL + l) * 64 + R.lane] = act && (f % fs == fs - 1) ? cabs2(v[l]) : 0.f;
      __syncthreads();
#pragma unroll
      for (int l = 0; l < CL; ++l) iest[l] = 0.f;
      for (int g = 0; g < nF; ++g) {
#pragma unroll
        for (int l = 0; l < CL; ++l) iest[l] += xs[((lslot + g) * CL + l) * 64 + R.lane];
      }
    }
#pragma unroll
    for (int l = 0; l < CL; ++l) {
      const int n = cslot<NP, CWR>(l, rho);
      const float im = ikt[q * csi + n];
      if (ITEM_LEAD(act, frw)) {
        err += (double)fabsf(im - iest[l]);
        // element (n, p) is detector pixel (brev(n), brev(p)) (the frames'
        // plan-time permutation)
        if (de) de[(size_t)brev(n, Line<NP>::LOG2) * NP + brev(p, Line<NP>::LOG2)] = fabsf(im - iest[l]);
      }
      v[l] = cscale(v[l], sqrtf(im / (iest[l] + 1e-10f)));
    }
    fft_inv_colg<NP, CWR, true>(R, v, rho);
    if (act) {
#pragma unroll
      for (int l = 0; l < CL; ++l) tile[frw * TS + q * cs + l * LPC + rho] = v[l];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < CL; ++i) {
      const int e = i * R.nt + (int)threadIdx.x;
      const int s = e >> lgNC, n = (e >> lgC) & (NP - 1), c = e & (C - 1);
      if (s < nF && c < Cv) cstore(R, W + s * Np2 + (size_t)n * NP + pc0 + c, tile[s * TS + c * cs + n]);
    }
    CITEM_ROUND_SYNC();
  CITEMS_END
  if (MX && R.wave == 0) {
    float m0, m1;
    max_finish(R, slot, mxl, m0, m1);
    if (R.lane == 0) {
      R.smx[3] = __float_as_uint(m0);
      R.smx[4] = __float_as_uint(m1);
    }
  }
  const double s = block_sum(R, err);
  if (threadIdx.x == 0) cstored(R, ep + (size_t)j * R.G + R.g, s);
}

// MX > 0: also read the position's maxima into LDS for the update pass, MX
// partial loads per lane and word (the staged form; the direct form leaves
// them to the update pass, MX = 0)
template <int NP, int CWR, int MX = 0>
__device__ __forceinline__ void detector_cols_fraunhofer(Run& R, int k, int j, double* ep, int fs, float* de,
                                                         const unsigned* slot = nullptr) {
  if constexpr (col_staged<NP>())
    detector_cols_fraunhofer_staged<NP, CWR, MX>(R, k, j, ep, fs, de, slot);
  else
    detector_cols_fraunhofer_direct<NP, CWR>(R, k, j, ep, fs, de);
}


// Column pass of an angular-spectrum propagation: forward column transform,
// multiply by the (bit-reversed, 1/Np^2-scaled) transfer function or its
// conjugate, inverse column transform.  Fields are independent, so the
// items are (field, column group) pairs and one wave takes one item.  A
// column group outside the transfer function's band limit is zero after the
// product; it is neither read nor written (the next row pass zeroes it), so
// only the live groups of each field are items (DevInst::lgrp), split into G
// contiguous ranges whose sizes differ by at most one: the live groups of a
// band-limited function lie scattered in the bit-reversed column order, and
// an equal share of them per workgroup keeps the slowest workgroup, which
// the barrier after the pass waits for, from carrying several.
// MX > 0: wave 0 also reads words 0, 1 of the maxima slot mslot for the pass
// after this one (MX partial loads per lane and word) and leaves them in LDS
// (smx[3], smx[4]), as the detector column passes do.
template <int NP, int CWR, bool PK, int MX = 0>
__device__ __forceinline__ void tf_cols(Run& R, bool conj, const unsigned* mslot = nullptr) {
  constexpr int CL = Col<NP, CWR>::CL, LPC = Col<NP, CWR>::LPC, NG = Col<NP, CWR>::NG;
  MaxLoads<MX ? MX : 1> mxl;
  if (MX && R.wave == 0) max_issue(R, mslot, mxl);
  const DevInst& D = *R.D;
  const size_t Np2 = (size_t)NP * NP;
  float2* __restrict__ W = D.W;
  const int rho = col_rho<NP, CWR>(R);
  const int* __restrict__ nl = D.nlive;
  const int ngl = nl ? nl[0] : NG;  // items per field
  const int T = R.nF * ngl;
  const int i0 = (int)(((long long)R.g * T) / R.G), i1 = (int)(((long long)(R.g + 1) * T) / R.G);
  for (int i = i0 + R.wave; i < i1; i += R.nwave) {
    const int f = i / ngl, m = i - f * ngl;
    const int tf = f / D.fpl;
    int cg = m;
    if (nl) {
      if (m >= nl[1 + tf]) continue;
      cg = D.lgrp[tf * NG + m];
    }
    const int p = cg * Col<NP, CWR>::CW + R.lane % Col<NP, CWR>::CW;
    // the transfer function of field f's wavelength
    const float2* __restrict__ HT = D.HT + (size_t)tf * Np2;
    // packed exchange rows: the m-th live group sits at columns m * CW
    const int ws = PK ? xrow_stride<NP>(D) : NP;
    const int pw = PK && D.cix ? m * Col<NP, CWR>::CW + R.lane % Col<NP, CWR>::CW : p;
    float2* const Wf = W + (size_t)f * NP * ws;
    float2 v[CL];
    load_colg<NP, CWR>(R, v, Wf, pw, rho, ws);
    fft_fwd_colg<NP, CWR, true>(R, v, rho);
#pragma unroll
    for (int l = 0; l < CL; ++l) {
      float2 h = ldg2(HT + (size_t)cslot<NP, CWR>(l, rho) * NP + p);
      if (conj) h = conjf2(h);
      v[l] = cmul(v[l], h);
    }
    fft_inv_colg<NP, CWR, true>(R, v, rho);
    store_colg<NP, CWR>(R, Wf, v, pw, rho, ws);
  }
  if (MX && R.wave == 0) {
    float m0, m1;
    max_finish(R, mslot, mxl, m0, m1);
    if (R.lane == 0) {
      R.smx[3] = __float_as_uint(m0);
      R.smx[4] = __float_as_uint(m1);
    }
  }
}

// ASP detector row pass: inverse row transform gives the detector field, the
// intensity constraint is applied in natural order, then the forward row
// transform starts the back-propagation.
template <int NP, bool PK, int MX = 0>
__device__ __forceinline__ void asp_rows_constraint(Run& R, int k, int j, double* ep, float* de,
                                                    const unsigned* slot = nullptr) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const size_t Np2 = (size_t)NP * NP;
  float2* __restrict__ W = D.W;
  const float* __restrict__ Ik = D.Ik + (size_t)k * Np2;
  const int nF = D.nF, fs = D.fs;
  const int ws = PK ? xrow_stride<NP>(D) : NP;
  // MX: wave 0 reads the position's maxima here, two passes before the
  // update pass needs them (they are final since the first barrier), with
  // its first item's loads in flight, and leaves them in LDS (smx[3], smx[4])
  MaxLoads<MX ? MX : 1> mxl;
  if (MX && R.wave == 0) max_issue(R, slot, mxl);
  double err = 0.0;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    float2 v[L];
    float im[L];
    const short* ix = PK ? packed_ix(R, D, f) : nullptr;
    const unsigned lb = ix ? ~0u : live_bits<NP>(D, f, R.t);
    float2* const Wr = W + (size_t)f * NP * ws + (size_t)r * ws;
    load_row_live<NP>(R, v, Wr, ix, lb, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) im[l] = ldg(Ik + (size_t)r * NP + l * LPF + R.t);
    fft_inv<NP, true>(v, R.tw, R.t, R.fx);
    float2 q[L];
#pragma unroll
    for (int l = 0; l < L; ++l) q[l] = make_float2(act && (f % fs == fs - 1) ? cabs2(v[l]) : 0.f, 0.f);
    lds_put<NP>(R.xa, islot, q, R.t);
    __syncthreads();
    float iest[L];
#pragma unroll
    for (int l = 0; l < L; ++l) iest[l] = 0.f;
    for (int g = 0; g < nF; ++g) {
#pragma unroll
      for (int l = 0; l < L; ++l) iest[l] += R.xa[(lslot + g) * NP + l * LPF + R.t].x;
    }
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if (ITEM_LEAD(act, frw)) {
        err += (double)fabsf(im[l] - iest[l]);
        if (de) de[(size_t)r * NP + l * LPF + R.t] = fabsf(im[l] - iest[l]);
      }
      v[l] = cscale(v[l], sqrtf(im[l] / (iest[l] + 1e-10f)));
    }
    fft_fwd<NP, true>(v, R.tw, R.t, R.fx);
    if (act) store_row_live<NP>(R, Wr, v, ix, R.t);
    ITEM_ROUND_SYNC();
  ITEMS_END
  if (MX && R.wave == 0) {
    float m0, m1;
    max_finish(R, slot, mxl, m0, m1);
    if (R.lane == 0) {
      R.smx[3] = __float_as_uint(m0);
      R.smx[4] = __float_as_uint(m1);
    }
  }
  const double s = block_sum(R, err);
  if (threadIdx.x == 0) cstored(R, ep + (size_t)j * R.G + R.g, s);
}

// Last row pass: inverse row transform gives the updated exit wave of the
// field, the difference to the exit wave of the same position, and the
// field's terms of the engine's object and probe increments (both from the
// pre-update patch and probe); the terms are summed over fields through LDS
// and each object / probe field is written by one owner slot (the field with
// psm index 0, resp. osm index 0).  Each object pixel of the patch has one
// owner.
template <int NP, bool QN, bool ASPF, bool QPH, bool PK, bool KEEP = false, bool MXL = false>
__device__ __forceinline__ void pie_rows_update(Run& R, int k, const unsigned* slot, bool sync_tail,
                                                const float2 (&ko)[Line<NP>::L], const float2 (&kp)[Line<NP>::L]) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int no = D.a.nosm, npm = D.a.npsm, ns = D.a.nslice;
  const int No = D.a.No, nF = D.nF;
  const size_t No2 = (size_t)No * No, Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  float2* __restrict__ O = (float2*)D.a.object;
  float2* __restrict__ P = (float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  const int prop = D.a.propagator;
  const bool cb = prop_checkerboard(prop);
  const float2* __restrict__ Qp = D.Q;
  const float invN = 1.0f / (float)NP;
  const int engine = D.a.engine;
  float Pmax, Omax;
  if (MXL) {
    // left in LDS by the constraint pass of this position
    Pmax = __uint_as_float(R.smx[3]);
    Omax = __uint_as_float(R.smx[4]);
  } else {
    load_max_published2(R, slot, Pmax, Omax);
  }
  const float bO = D.a.betaObject;
  const float stepP = engine == PTY_MPIE ? D.a.probeWeight * D.a.betaProbe : D.a.betaProbe;
  // ePIE.py:128-160 weights conj(P)/max(sum|P|^2), conj(O)/max(sum|O|^2);
  // mPIE.py:226-270 / multiPIE.py:153-205 (CPM) weights
  // conj(P)/(aO Pmax + (1-aO)|P|^2), conj(O)/(aP Omax + (1-aP)|O|^2)
  const float invP = 1.0f / Pmax, invO = 1.0f / Omax;
  const float aPm = D.a.alphaObject * Pmax, aOm = D.a.alphaProbe * Omax;
  const float omO = D.omAlphaO, omP = D.omAlphaP;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    (void)frw;
    const FieldIdx x = field_idx(f, no, npm, ns);
    const unsigned osel = D.osel[f], psel = D.psel[f];
    const size_t orow_off = (size_t)(row0 + r) * No + col0;
    const size_t prow_off = (size_t)r * NP;
    float2 d[L], of[L], pf[L];
    const short* ix = ASPF && PK ? packed_ix(R, D, f) : nullptr;
    if (ix) {
      // the packed frequency-domain row after the patch and probe rows,
      // whose loads need no column table
      const int ws = PK ? xrow_stride<NP>(D) : NP;
      if constexpr (KEEP) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
          of[l] = ko[l];
          pf[l] = kp[l];
        }
      } else {
        load_row<NP>(R, of, O + x.fo * No2 + orow_off, R.t);
        load_row<NP>(R, pf, P + x.fp * Np2 + prow_off, R.t);
      }
      load_row_live<NP>(R, d, W + (size_t)f * NP * ws + (size_t)r * ws, ix, ~0u, R.t);
    } else if constexpr (ASPF) {
      // the column flags are read before the row loads, whose latency hides theirs
      const unsigned lb = live_bits<NP>(D, f, R.t);
      load_row_f<NP>(R, d, W + f * Np2 + prow_off, R.t);
      if constexpr (KEEP) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
          of[l] = ko[l];
          pf[l] = kp[l];
        }
      } else {
        load_row<NP>(R, of, O + x.fo * No2 + orow_off, R.t);
        load_row<NP>(R, pf, P + x.fp * Np2 + prow_off, R.t);
      }
      zero_dead_columns<NP>(lb, d);
    } else {
      load_row_f<NP>(R, d, W + f * Np2 + prow_off, R.t);
      if constexpr (KEEP) {
#pragma unroll
        for (int l = 0; l < L; ++l) {
          of[l] = ko[l];
          pf[l] = kp[l];
        }
      } else {
        load_row<NP>(R, of, O + x.fo * No2 + orow_off, R.t);
        load_row<NP>(R, pf, P + x.fp * Np2 + prow_off, R.t);
      }
    }
    fft_inv<NP, true>(d, R.tw, R.t, R.fx);
    float2 tO[L], tP[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float2 u = d[l];
      if constexpr (!ASPF) {
        u = cscale(u, invN);
        if (cb && ((r + l * LPF + R.t) & 1)) u = make_float2(-u.x, -u.y);
        // Fresnel: ifft2c(.) * conj(Q) (Operators.py:150)
        if (Qp) u = cmul(u, conjf2(ldg2(Qp + prow_off + l * LPF + R.t)));
      } else if constexpr (QPH) {
        // scaled ASP: ... conj(c Q1) (Operators.py:343-344), per wavelength
        u = cmul(u, conjf2(ldg2(Qp + (size_t)(f / D.fpl) * Np2 + prow_off + l * LPF + R.t)));
      }
      const float2 dl = csub(u, cmul(of[l], pf[l]));
      float wo, wp;
      if constexpr (QN) {
        // qNewton.py:116-121: (|P| / Pmax) conj(P) / (|P|^2 + regObject); the
        // probe step waits for Omax over the updated object (qn_probe_rows)
        const float ap = sqrtf(cabs2(pf[l]));
        const float a = ap / Pmax;
        const float2 b = make_float2(a * pf[l].x, -(a * pf[l].y));
        const float sc = 1.0f / (cabs2(pf[l]) + D.a.alphaObject);
        tO[l] = cmul(make_float2(b.x * sc, b.y * sc), dl);
        tP[l] = make_float2(0.f, 0.f);
        d[l] = dl;
        continue;
      }
      if (engine == PTY_EPIE) {
        wo = invP;
        wp = invO;
      } else {
        wo = 1.0f / (aPm + omO * cabs2(pf[l]));
        wp = 1.0f / (aOm + omP * cabs2(of[l]));
      }
      tO[l] = cmul(make_float2(pf[l].x * wo, -pf[l].y * wo), dl);
      tP[l] = cmul(make_float2(of[l].x * wp, -of[l].y * wp), dl);
    }
    constexpr bool qn = QN;
    if (!qn && nF == 1 && (osel & psel & 1u)) {
      // one field: its own terms are the sums of the LDS form (0 + x), no
      // exchange and no workgroup barrier
      if (act) {
        float2 nv[L];
#pragma unroll
        for (int l = 0; l < L; ++l) nv[l] = cadd(of[l], cscale(cadd(make_float2(0.f, 0.f), tO[l]), bO));
        if (D.pcpart) store_row<NP>(R, D.OP + x.fo * Np2 + prow_off, of, R.t);
        store_row<NP>(R, O + x.fo * No2 + orow_off, nv, R.t);
#pragma unroll
        for (int l = 0; l < L; ++l)
          nv[l] = cadd(pf[l], cscale(cadd(make_float2(0.f, 0.f), tP[l]), engine == PTY_EPIE ? D.a.betaProbe : stepP));
        store_row<NP>(R, P + x.fp * Np2 + prow_off, nv, R.t);
      }
      continue;
    }
    lds_put<NP>(R.xa, islot, tO, R.t);
    // a probe field that takes its own term alone (mPIE with one object
    // mode: the sum runs over osm) is written at once, without the exchange
    const bool pown = !qn && psel == (1u << f);
    if (pown) {
      if (act && x.io == 0) {
        float2 nv[L];
#pragma unroll
        for (int l = 0; l < L; ++l)
          nv[l] = cadd(pf[l], cscale(cadd(make_float2(0.f, 0.f), tP[l]), engine == PTY_EPIE ? D.a.betaProbe : stepP));
        store_row<NP>(R, P + x.fp * Np2 + prow_off, nv, R.t);
      }
    } else {
      lds_put<NP>(R.xb, islot, tP, R.t);
    }
    // qNewton keeps the difference for its probe step; only this item
    // touches row r of field f in this pass
    if (qn && act) store_row<NP>(R, W + f * Np2 + prow_off, d, R.t);
    __syncthreads();
    if (act && x.ip == 0) {
      // object field x.fo: ePIE and qNewton sum over (lambda, psm, slice),
      // the m/multi engines over psm
      float2 nv[L], acc[L];
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = make_float2(0.f, 0.f);
      for (int g = 0; g < nF; ++g) {
        if ((osel >> g) & 1u) {
#pragma unroll
          for (int l = 0; l < L; ++l) acc[l] = cadd(acc[l], R.xa[(lslot + g) * NP + l * LPF + R.t]);
        }
      }
#pragma unroll
      for (int l = 0; l < L; ++l) nv[l] = cadd(of[l], cscale(acc[l], bO));
      // the pre-update patch: qNewton's probe step and the position correction read it
      if (qn || D.pcpart) store_row<NP>(R, D.OP + x.fo * Np2 + prow_off, of, R.t);
      store_row<NP>(R, O + x.fo * No2 + orow_off, nv, R.t);
    }
    if (act && x.io == 0 && !qn && !pown) {
      // probe field x.fp: ePIE sums over (lambda, osm, slice), mPIE over osm,
      // multiPIE over (lambda, osm)
      float2 nv[L], acc[L];
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = make_float2(0.f, 0.f);
      for (int g = 0; g < nF; ++g) {
        if ((psel >> g) & 1u) {
#pragma unroll
          for (int l = 0; l < L; ++l) acc[l] = cadd(acc[l], R.xb[(lslot + g) * NP + l * LPF + R.t]);
        }
      }
#pragma unroll
      for (int l = 0; l < L; ++l) nv[l] = cadd(pf[l], cscale(acc[l], engine == PTY_EPIE ? D.a.betaProbe : stepP));
      store_row<NP>(R, P + x.fp * Np2 + prow_off, nv, R.t);
    }
    // the next position's first pass rewrites the LDS slots at once when the
    // cluster barrier after this pass is skipped
    if (sync_tail) __syncthreads(); else ITEM_ROUND_SYNC();
  ITEMS_END
}

// Field-loop forms of the first and last row passes, for workgroups that own
// at least one line per item slot (clusters of a few workgroups): one slot
// takes a line and loops over its fields, so the field sums stay in
// registers, in increasing field order as in the LDS forms (bit-identical
// results), and a pass has no workgroup barrier.  Only for one object field
// whose owner sums every field and probe field f taking field f alone
// (ePIE / mPIE / multiPIE with nlambda = nosm = nslice = 1, DevInst::fl).
template <int NP, bool ASPF, bool QPH>
__device__ __forceinline__ void pie_rows_first_fl(Run& R, int k, unsigned* slot, unsigned* zslot) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int No = D.a.No, nF = D.nF;
  const size_t Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  const float2* __restrict__ O = (const float2*)D.a.object;
  const float2* __restrict__ P = (const float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  const bool cb = prop_checkerboard(D.a.propagator);
  const float2* __restrict__ Qp = D.Q;
  (void)zslot;
  float pmax = 0.f, omax = 0.f;
  const int per = NP / R.G, line0 = R.g * per;
  const int islot = R.wave * Line<NP>::SUB + R.sub;
  for (int base = 0; base < per; base += R.ipr) {
    const int li = base + islot;
    const bool act = li < per;
    const int r = line0 + (act ? li : per - 1);
    float2 of[L];
    load_row<NP>(R, of, O + (size_t)(row0 + r) * No + col0, R.t);
    float sp[L], so[L];
#pragma unroll
    for (int l = 0; l < L; ++l) sp[l] = so[l] = 0.f;
    for (int f = 0; f < nF; ++f) {
      float2 pf[L], v[L];
      load_row<NP>(R, pf, P + f * Np2 + (size_t)r * NP, R.t);
#pragma unroll
      for (int l = 0; l < L; ++l) {
        sp[l] += act ? cabs2(pf[l]) : 0.f;
        so[l] += act && f == 0 ? cabs2(of[l]) : 0.f;
        float2 e = cmul(of[l], pf[l]);
        if constexpr (!ASPF || QPH) {
          if (Qp) e = cmul(e, ldg2(Qp + (size_t)(f / D.fpl) * Np2 + (size_t)r * NP + l * LPF + R.t));
        }
        if (cb && ((r + l * LPF + R.t) & 1)) e = make_float2(-e.x, -e.y);
        v[l] = e;
      }
      fft_fwd<NP, true>(v, R.tw, R.t, R.fx);
      if constexpr (ASPF) {
        const int ws = xrow_stride<NP>(D);
        if (act) store_row_live<NP>(R, W + (size_t)f * NP * ws + (size_t)r * ws, v, packed_ix(R, D, f), R.t);
      } else {
        if (act) store_row_f<NP>(R, W + f * Np2 + (size_t)r * NP, v, R.t);
      }
    }
    if (act) {
#pragma unroll
      for (int l = 0; l < L; ++l) {
        pmax = fmaxf(pmax, sp[l]);
        omax = fmaxf(omax, so[l]);
      }
    }
  }
  publish_max2(R, slot, pmax, omax);
}

template <int NP, bool ASPF, bool QPH>
__device__ __forceinline__ void pie_rows_update_fl(Run& R, int k, const unsigned* slot) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int No = D.a.No, nF = D.nF;
  const size_t Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  float2* __restrict__ O = (float2*)D.a.object;
  float2* __restrict__ P = (float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  const bool cb = prop_checkerboard(D.a.propagator);
  const float2* __restrict__ Qp = D.Q;
  const float invN = 1.0f / (float)NP;
  const int engine = D.a.engine;
  float Pmax, Omax;
  load_max_published2(R, slot, Pmax, Omax);
  const float bO = D.a.betaObject;
  const float stepP = engine == PTY_MPIE ? D.a.probeWeight * D.a.betaProbe : D.a.betaProbe;
  const float bP = engine == PTY_EPIE ? D.a.betaProbe : stepP;
  const float invP = 1.0f / Pmax, invO = 1.0f / Omax;
  const float aPm = D.a.alphaObject * Pmax, aOm = D.a.alphaProbe * Omax;
  const float omO = D.omAlphaO, omP = D.omAlphaP;
  const int per = NP / R.G, line0 = R.g * per;
  const int islot = R.wave * Line<NP>::SUB + R.sub;
  for (int base = 0; base < per; base += R.ipr) {
    const int li = base + islot;
    const bool act = li < per;
    const int r = line0 + (act ? li : per - 1);
    const size_t orow_off = (size_t)(row0 + r) * No + col0;
    const size_t prow_off = (size_t)r * NP;
    float2 of[L], accO[L];
    load_row<NP>(R, of, O + orow_off, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) accO[l] = make_float2(0.f, 0.f);
    for (int f = 0; f < nF; ++f) {
      float2 d[L], pf[L], nv[L];
      const short* ix = ASPF ? packed_ix(R, D, f) : nullptr;
      if (ix) {
        const int ws = xrow_stride<NP>(D);
        load_row<NP>(R, pf, P + f * Np2 + prow_off, R.t);
        load_row_live<NP>(R, d, W + (size_t)f * NP * ws + (size_t)r * ws, ix, ~0u, R.t);
      } else if constexpr (ASPF) {
        const unsigned lb = live_bits<NP>(D, f, R.t);
        load_row_f<NP>(R, d, W + f * Np2 + prow_off, R.t);
        load_row<NP>(R, pf, P + f * Np2 + prow_off, R.t);
        zero_dead_columns<NP>(lb, d);
      } else {
        load_row_f<NP>(R, d, W + f * Np2 + prow_off, R.t);
        load_row<NP>(R, pf, P + f * Np2 + prow_off, R.t);
      }
      fft_inv<NP, true>(d, R.tw, R.t, R.fx);
#pragma unroll
      for (int l = 0; l < L; ++l) {
        float2 u = d[l];
        if constexpr (!ASPF) {
          u = cscale(u, invN);
          if (cb && ((r + l * LPF + R.t) & 1)) u = make_float2(-u.x, -u.y);
          if (Qp) u = cmul(u, conjf2(ldg2(Qp + prow_off + l * LPF + R.t)));
        } else if constexpr (QPH) {
          u = cmul(u, conjf2(ldg2(Qp + (size_t)(f / D.fpl) * Np2 + prow_off + l * LPF + R.t)));
        }
        const float2 dl = csub(u, cmul(of[l], pf[l]));
        float wo, wp;
        if (engine == PTY_EPIE) {
          wo = invP;
          wp = invO;
        } else {
          wo = 1.0f / (aPm + omO * cabs2(pf[l]));
          wp = 1.0f / (aOm + omP * cabs2(of[l]));
        }
        accO[l] = cadd(accO[l], cmul(make_float2(pf[l].x * wo, -pf[l].y * wo), dl));
        const float2 accP = cadd(make_float2(0.f, 0.f), cmul(make_float2(of[l].x * wp, -of[l].y * wp), dl));
        nv[l] = cadd(pf[l], cscale(accP, bP));
      }
      if (act) store_row<NP>(R, P + f * Np2 + prow_off, nv, R.t);
    }
    if (act) {
      float2 nv[L];
#pragma unroll
      for (int l = 0; l < L; ++l) nv[l] = cadd(of[l], cscale(accO[l], bO));
      store_row<NP>(R, O + orow_off, nv, R.t);
    }
  }
}

// qNewton, after the object step: Omax = max over the whole object of the
// sum over its fields of |O| (qNewton.py:134), each workgroup over its own
// chunk of pixels, fields added in C order.
template <int NP>
__device__ __forceinline__ void qn_omax_pass(Run& R, unsigned* slot) {
  const DevInst& D = *R.D;
  const size_t n = (size_t)D.a.No * D.a.No;
  const float2* __restrict__ O = (const float2*)D.a.object;
  const size_t chunk = (n + R.G - 1) / R.G;
  const size_t b = chunk * R.g;
  const size_t e = b + chunk < n ? b + chunk : n;
  float m = 0.f;
  for (size_t i = b + threadIdx.x; i < e; i += R.nt) {
    float s = 0.f;
    for (int fo = 0; fo < D.nO; ++fo) s += sqrtf(cabs2(cload(R, O + fo * n + i)));
    m = fmaxf(m, s);
  }
  m = wave_max(m);
  if (R.lane == 0) atomic_fmax_slot(R, slot, SLOT_QN_OMAX, m);
}

// qNewton probe step (qNewton.py:129-140): probe += betaProbe sum over
// (lambda, osm, slice) of (|O|/Omax) conj(O) / (|O|^2 + regProbe) * DELTA,
// O the patch before this position's object step.
template <int NP>
__device__ __forceinline__ void qn_probe_rows(Run& R, const unsigned* slot) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int no = D.a.nosm, npm = D.a.npsm, ns = D.a.nslice;
  const int nF = D.nF;
  const size_t Np2 = (size_t)NP * NP;
  float2* __restrict__ P = (float2*)D.a.probe;
  const float Omax = load_slot(R, slot, SLOT_QN_OMAX);
  const float reg = D.a.alphaProbe, bP = D.a.betaProbe;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    (void)frw;
    const FieldIdx x = field_idx(f, no, npm, ns);
    const unsigned psel = D.psel[f];
    const size_t prow_off = (size_t)r * NP;
    float2 d[L], op[L], t[L];
    load_row<NP>(R, d, D.W + f * Np2 + prow_off, R.t);
    load_row<NP>(R, op, D.OP + x.fo * Np2 + prow_off, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const float a = sqrtf(cabs2(op[l])) / Omax;
      const float2 b = make_float2(a * op[l].x, -(a * op[l].y));
      const float sc = 1.0f / (cabs2(op[l]) + reg);
      t[l] = cmul(make_float2(b.x * sc, b.y * sc), d[l]);
    }
    lds_put<NP>(R.xb, islot, t, R.t);
    __syncthreads();
    if (act && x.io == 0) {
      float2 pf[L], nv[L];
      load_row<NP>(R, pf, P + x.fp * Np2 + prow_off, R.t);
      float2 acc[L];
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = make_float2(0.f, 0.f);
      for (int g = 0; g < nF; ++g) {
        if ((psel >> g) & 1u) {
#pragma unroll
          for (int l = 0; l < L; ++l) acc[l] = cadd(acc[l], R.xb[(lslot + g) * NP + l * LPF + R.t]);
        }
      }
#pragma unroll
      for (int l = 0; l < L; ++l) nv[l] = cadd(pf[l], cscale(acc[l], bP));
      store_row<NP>(R, P + x.fp * Np2 + prow_off, nv, R.t);
    }
    __syncthreads();
  ITEMS_END
}

// mPIE / multiPIE momentum step over the whole object and probe
// (mPIE.py:196-224, multiPIE.py:125-151); each workgroup owns a contiguous
// chunk, so the step needs no read-modify-write atomics.
template <int NP>
__device__ __forceinline__ void momentum_pass(Run& R) {
  const DevInst& D = *R.D;
  const float eta = D.a.momFriction, nu = D.a.momFeedback;
  for (int which = 0; which < 2; ++which) {
    const size_t n = which == 0 ? (size_t)D.nO * D.a.No * D.a.No : (size_t)D.nPr * D.a.Np * D.a.Np;
    float2* X = (float2*)(which == 0 ? D.a.object : D.a.probe);
    float2* B = (float2*)(which == 0 ? D.a.objectBuffer : D.a.probeBuffer);
    float2* M = (float2*)(which == 0 ? D.a.objectMomentum : D.a.probeMomentum);
    const size_t chunk = (n + R.G - 1) / R.G;
    const size_t b = chunk * R.g;
    const size_t e = b + chunk < n ? b + chunk : n;
    // MU elements per thread at a time, all their loads issued before the
    // first store (a store can alias the next element's loads, so the
    // compiler would otherwise wait out one round trip per element)
    constexpr int MU = 4;
    size_t i = b + threadIdx.x;
    for (; i + (size_t)(MU - 1) * R.nt < e; i += (size_t)MU * R.nt) {
      float2 x[MU], bb[MU], m[MU];
#pragma unroll
      for (int u = 0; u < MU; ++u) {
        const size_t iu = i + (size_t)u * R.nt;
        x[u] = cload(R, X + iu);
        bb[u] = cload(R, B + iu);
        m[u] = cload(R, M + iu);
      }
#pragma unroll
      for (int u = 0; u < MU; ++u) {
        const size_t iu = i + (size_t)u * R.nt;
        const float2 mn = make_float2((bb[u].x - x[u].x) + m[u].x * eta, (bb[u].y - x[u].y) + m[u].y * eta);
        const float2 xn = make_float2(x[u].x - mn.x * nu, x[u].y - mn.y * nu);
        cstore(R, M + iu, mn);
        cstore(R, X + iu, xn);
        cstore(R, B + iu, xn);
      }
    }
    for (; i < e; i += R.nt) {
      const float2 x = cload(R, X + i), bb = cload(R, B + i), m = cload(R, M + i);
      const float2 mn = make_float2((bb.x - x.x) + m.x * eta, (bb.y - x.y) + m.y * eta);
      const float2 xn = make_float2(x.x - mn.x * nu, x.y - mn.y * nu);
      cstore(R, M + i, mn);
      cstore(R, X + i, xn);
      cstore(R, B + i, xn);
    }
  }
}

// mqNewton's momentum step after every position (mqNewton.py:121-122,
// 184-224), on the gradient g = buffer - current of the whole object and the
// whole probe, with itr = iteration + 1:
//   ADAM (142-153): m = b1 m + (1 - b1) g, v = b2 v + (1 - b2) |g|^2 (the
//     squared 2-norm of the whole gradient, one number), step =
//     (m / (1 - b1^itr)) / (sqrt(v / (1 - b2^itr)) + 1e-8);
//   NADAM (155-174): the same m and v, step = (b1 m / (1 - b1^itr) +
//     g (1 - b1) / (1 - b1^itr)) / (sqrt(v / (1 - b2^itr)) + 1e-8);
//   momentum (176-182): m = g + friction m, step = m (v untouched);
// then current -= beta_m step, buffer = current.  ADAM and NADAM take two
// passes: FP64 partials of |g|^2 per workgroup, then (after a barrier) every
// workgroup sums them in workgroup order and updates its own chunk; the
// momentum method needs only the second.  The instance fields carry b1 (or
// the friction) in momFriction, b2 in momFeedback and the object and probe
// steps beta_m in probeWeight and sliceBeta; the method is PTY_FLAG_MQN_*.
__device__ __forceinline__ void adam_norm_pass(Run& R) {
  const DevInst& D = *R.D;
  for (int which = 0; which < 2; ++which) {
    const size_t n = which == 0 ? (size_t)D.nO * D.a.No * D.a.No : (size_t)D.nPr * D.a.Np * D.a.Np;
    const float2* X = (const float2*)(which == 0 ? D.a.object : D.a.probe);
    const float2* B = (const float2*)(which == 0 ? D.a.objectBuffer : D.a.probeBuffer);
    const size_t chunk = (n + R.G - 1) / R.G;
    const size_t b = chunk * R.g;
    const size_t e = b + chunk < n ? b + chunk : n;
    double acc = 0.0;
    for (size_t i = b + threadIdx.x; i < e; i += R.nt) {
      const float2 x = cload(R, X + i), bb = cload(R, B + i);
      const float gx = bb.x - x.x, gy = bb.y - x.y;
      acc += (double)gx * (double)gx + (double)gy * (double)gy;
    }
    const double s = block_sum(R, acc);
    if (threadIdx.x == 0) cstored(R, D.adampart + (size_t)which * R.G + R.g, s);
  }
}

__device__ __forceinline__ void adam_update_pass(Run& R, int itr, unsigned long long p) {
  const DevInst& D = *R.D;
  const bool nadam = (D.a.flags & PTY_FLAG_MQN_NADAM) != 0;
  const bool plain = (D.a.flags & PTY_FLAG_MQN_MOMENTUM) != 0;
  const float b1 = D.a.momFriction, b2 = D.a.momFeedback;
  const float omb1 = (float)(1.0 - (double)b1), omb2 = (float)(1.0 - (double)b2);
  const float s1 = (float)(1.0 - pow((double)b1, (double)itr));
  const float s2 = (float)(1.0 - pow((double)b2, (double)itr));
  for (int which = 0; which < 2; ++which) {
    float den = 1.f;
    if (!plain) {
      double sq = 0.0;
      for (int g = 0; g < R.G; ++g) sq += cloadd(R, D.adampart + (size_t)which * R.G + g);
      const float nrm = sqrtf((float)sq);
      const float vt = b2 * (float)cloadd(R, D.adamv + (p & 1ull) * 2 + which) + omb2 * (nrm * nrm);
      if (R.g == 0 && threadIdx.x == 0) cstored(R, D.adamv + ((p + 1ull) & 1ull) * 2 + which, (double)vt);
      den = sqrtf(vt / s2) + 1e-8f;
    } else if (R.g == 0 && threadIdx.x == 0) {
      // the second moment passes through unchanged
      cstored(R, D.adamv + ((p + 1ull) & 1ull) * 2 + which, cloadd(R, D.adamv + (p & 1ull) * 2 + which));
    }
    const float bm = which == 0 ? D.a.probeWeight : D.a.sliceBeta;
    const size_t n = which == 0 ? (size_t)D.nO * D.a.No * D.a.No : (size_t)D.nPr * D.a.Np * D.a.Np;
    float2* X = (float2*)(which == 0 ? D.a.object : D.a.probe);
    float2* B = (float2*)(which == 0 ? D.a.objectBuffer : D.a.probeBuffer);
    float2* M = (float2*)(which == 0 ? D.a.objectMomentum : D.a.probeMomentum);
    const size_t chunk = (n + R.G - 1) / R.G;
    const size_t b = chunk * R.g;
    const size_t e = b + chunk < n ? b + chunk : n;
    for (size_t i = b + threadIdx.x; i < e; i += R.nt) {
      const float2 x = cload(R, X + i), bb = cload(R, B + i), m = cload(R, M + i);
      const float2 gr = make_float2(bb.x - x.x, bb.y - x.y);
      float2 mn, u;
      if (plain) {
        mn = make_float2(gr.x + b1 * m.x, gr.y + b1 * m.y);
        u = mn;
      } else {
        mn = make_float2(b1 * m.x + omb1 * gr.x, b1 * m.y + omb1 * gr.y);
        const float2 mh = make_float2(mn.x / s1, mn.y / s1);
        if (nadam) {
          const float2 gs = make_float2((gr.x * omb1) / s1, (gr.y * omb1) / s1);
          u = make_float2((b1 * mh.x + gs.x) / den, (b1 * mh.y + gs.y) / den);
        } else {
          u = make_float2(mh.x / den, mh.y / den);
        }
      }
      const float2 xn = make_float2(x.x - bm * u.x, x.y - bm * u.y);
      cstore(R, M + i, mn);
      cstore(R, X + i, xn);
      cstore(R, B + i, xn);
    }
  }
}

// Position correction (BaseEngine.positionCorrection, BaseEngine.py:1241-1323,
// the direct branch of positionCorrectionSwitch_radius < 2): cc_s =
// |Re sum conj(roll(Opatch, (rs, cs))) O[sy, sx]| (the reference stores the
// complex sum into a real array, :1256 and :1270-1272, which keeps the real
// part) for the 9 shifts rs, cs in
// {-1, 0, 1} (rowShifts / colShifts of :191-194), Opatch the patch before
// this position's update and O[sy, sx] the updated one; normFactor =
// sum |Opatch|^2.  Each workgroup sums its rows in FP64; the partials of a
// position are combined by workgroup 0 after a barrier (pc_direction_update).
template <int NP>
__device__ __forceinline__ void pc_corr_pass(Run& R, int k) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int No = D.a.No;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  const float2* O = (const float2*)D.a.object;
  const float2* OP = D.OP;
  double acc[PC_TERMS];
#pragma unroll
  for (int s = 0; s < PC_TERMS; ++s) acc[s] = 0.0;
  const int per = NP / R.G;
  const int line0 = R.g * per;
  for (int li = R.wave * Line<NP>::SUB + R.sub; li < per; li += R.nwave * Line<NP>::SUB) {
    const int i = line0 + li;
    float2 on[L];
    load_row<NP>(R, on, O + (size_t)(row0 + i) * No + col0, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int j = l * LPF + R.t;
#pragma unroll
      for (int s = 0; s < 9; ++s) {
        const int rs = s / 3 - 1, cs = s % 3 - 1;
        // roll(x, r)[i] = x[i - r] on the periodic patch
        const float2 v = cload(R, OP + (size_t)((i - rs) & (NP - 1)) * NP + ((j - cs) & (NP - 1)));
        // conj(v) * on
        acc[2 * s] += (double)v.x * on[l].x + (double)v.y * on[l].y;
        acc[2 * s + 1] += (double)v.x * on[l].y - (double)v.y * on[l].x;
        if (s == 4) acc[18] += (double)v.x * v.x + (double)v.y * v.y;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < PC_TERMS; ++s) {
    const double t = block_sum(R, acc[s]);
    if (threadIdx.x == 0) cstored(R, D.pcpart + (size_t)R.g * PC_TERMS + s, t);
  }
}

// D[k] = daleth clip(grad) + beth D[k] with daleth = 0.5, beth = 0.9,
// grad = betaGrad sum_s (cc_s - mean cc) / normFactor (colShift_s,
// rowShift_s), betaGrad = 1000, clipped at r = 3 pixels
// (BaseEngine.py:1305-1323, parameters :179-195); workgroup 0, thread 0.
__device__ __forceinline__ void pc_direction_update(Run& R, int k) {
  const DevInst& D = *R.D;
  if (R.g != 0 || threadIdx.x != 0) return;
  double t[PC_TERMS];
  for (int s = 0; s < PC_TERMS; ++s) t[s] = 0.0;
  for (int g = 0; g < R.G; ++g)
    for (int s = 0; s < PC_TERMS; ++s) t[s] += cloadd(R, D.pcpart + (size_t)g * PC_TERMS + s);
  double cc[9], mean = 0.0;
  for (int s = 0; s < 9; ++s) {
    cc[s] = fabs(t[2 * s]);
    mean += cc[s];
  }
  mean /= 9.0;
  double gx = 0.0, gy = 0.0;
  for (int s = 0; s < 9; ++s) {
    const double w = (cc[s] - mean) / t[18];
    gx += w * (double)(s % 3 - 1);
    gy += w * (double)(s / 3 - 1);
  }
  gx *= 1000.0;
  gy *= 1000.0;
  if (fabs(gx) > 3.0) gx = 3.0 * gx / fabs(gx);
  if (fabs(gy) > 3.0) gy = 3.0 * gy / fabs(gy);
  double* Dk = D.a.pc_direction + 2 * (size_t)k;
  Dk[0] = 0.5 * gy + 0.9 * Dk[0];
  Dk[1] = 0.5 * gx + 0.9 * Dk[1];
}

// Fourier branch of the position correction (BaseEngine.py:1277-1302,
// positionCorrectionSwitch_radius r >= 2): cc = |fftshift(ifft2(fft2(O[sy, sx]
// - mean) conj(fft2(Opatch - O.mean()))))| on the (2r + 1)^2 shifts around
// the centre, Opatch the patch before this position's update.  For the
// circular correlation the mean of the whole object drops out (the mean-free
// updated patch sums to zero), so the product's zero-frequency term is
// removed instead of subtracting the means.  Three passes over the two-field
// workspace PCW: forward row transforms of both patches (with the FP64
// partials of normFactor = sum |Opatch|^2), forward column transforms /
// product / inverse column transform keeping the 2r + 1 rows of the window,
// inverse row transforms of those rows giving |xcor| (numpy's ifft2 divides
// by Np^2); then workgroup 0 forms the search direction (pcf_direction_update).
template <int NP>
__device__ __forceinline__ void pcf_rows(Run& R, int k) {
  constexpr int L = Line<NP>::L;
  const DevInst& D = *R.D;
  const int No = D.a.No;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  const float2* O = (const float2*)D.a.object;
  const size_t Np2 = (size_t)NP * NP;
  double nrm = 0.0;
  const int per = NP / R.G;
  const int line0 = R.g * per;
  for (int li = R.wave * Line<NP>::SUB + R.sub; li < per; li += R.nwave * Line<NP>::SUB) {
    const int i = line0 + li;
    float2 a[L], b[L];
    load_row<NP>(R, a, O + (size_t)(row0 + i) * No + col0, R.t);
    load_row<NP>(R, b, D.OP + (size_t)i * NP, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) nrm += (double)b[l].x * b[l].x + (double)b[l].y * b[l].y;
    fft_fwd<NP>(a, R.tw, R.t, R.fx);
    fft_fwd<NP>(b, R.tw, R.t, R.fx);
    store_row<NP>(R, D.PCW + (size_t)i * NP, a, R.t);
    store_row<NP>(R, D.PCW + Np2 + (size_t)i * NP, b, R.t);
  }
  const double s = block_sum(R, nrm);
  if (threadIdx.x == 0) cstored(R, D.pcpart + (size_t)R.g * PC_TERMS + 18, s);
}

template <int NP, int CWR>
__device__ __forceinline__ void pcf_cols(Run& R) {
  constexpr int CL = Col<NP, CWR>::CL, LPC = Col<NP, CWR>::LPC, CW = Col<NP, CWR>::CW, NG = Col<NP, CWR>::NG;
  const DevInst& D = *R.D;
  const size_t Np2 = (size_t)NP * NP;
  const int r = D.a.pc_radius;
  const int rho = col_rho<NP, CWR>(R);
  const int cpg = (NG + R.G - 1) / R.G;
  const int cg0 = R.g * cpg;
  const int cgn = cg0 < NG ? (NG - cg0 < cpg ? NG - cg0 : cpg) : 0;
  for (int gi = R.wave; gi < cgn; gi += R.nwave) {
    const int p = (cg0 + gi) * CW + R.lane % CW;
    float2 a[CL], b[CL];
    load_colg<NP, CWR>(R, a, D.PCW, p, rho);
    load_colg<NP, CWR>(R, b, D.PCW + Np2, p, rho);
    fft_fwd_colg<NP, CWR>(R, a, rho);
    fft_fwd_colg<NP, CWR>(R, b, rho);
#pragma unroll
    for (int l = 0; l < CL; ++l) a[l] = cmul(a[l], conjf2(b[l]));
    // element (n, p) holds frequency (brev(n), brev(p)): zero frequency at n = p = 0
    if (p == 0 && rho == 0) a[0] = make_float2(0.f, 0.f);
    fft_inv_colg<NP, CWR>(R, a, rho);
#pragma unroll
    for (int l = 0; l < CL; ++l) {
      const int n = l * LPC + rho;  // natural row order after the inverse transform
      if (n <= r || n >= NP - r) cstore(R, D.PCW + (size_t)n * NP + p, a[l]);
    }
  }
}

template <int NP>
__device__ __forceinline__ void pcf_window_rows(Run& R) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF, SUB = Line<NP>::SUB;
  const DevInst& D = *R.D;
  const int r = D.a.pc_radius, nw = 2 * r + 1;
  const float inv = 1.0f / ((float)NP * (float)NP);
  const int step = R.G * R.nwave * SUB;
  for (int w = (R.g * R.nwave + R.wave) * SUB + R.sub; w < nw; w += step) {
    const int y = (w - r) & (NP - 1);  // row shift w - r
    float2 v[L];
    load_row<NP>(R, v, D.PCW + (size_t)y * NP, R.t);
    fft_inv<NP>(v, R.tw, R.t, R.fx);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const int x = l * LPF + R.t;
      const int dx = x <= r ? x : x - NP;  // column shift
      if (dx >= -r && dx <= r) cstored(R, D.pcc + (size_t)w * nw + (dx + r), (double)(sqrtf(cabs2(v[l])) * inv));
    }
  }
}

// D[k] = daleth clip(grad) + beth D[k] with grad = betaGrad sum_s (cc_s -
// mean cc) / normFactor (colShift_s, rowShift_s), betaGrad = 5, clipped at
// r = 10 pixels (BaseEngine.py:1295-1323); rowShifts / colShifts in np.mgrid
// order (row shift slowest).  Every thread of workgroup 0 takes a share of
// the shifts.
__device__ __forceinline__ void pcf_direction_update(Run& R, int k) {
  const DevInst& D = *R.D;
  if (R.g != 0) return;
  const int r = D.a.pc_radius, nw = 2 * r + 1, n = nw * nw;
  double sc = 0.0;
  for (int s = threadIdx.x; s < n; s += R.nt) sc += cloadd(R, D.pcc + s);
  sc = block_sum(R, sc);
  if (threadIdx.x == 0) R.red[0] = sc;
  __syncthreads();
  const double mean = R.red[0] / (double)n;
  __syncthreads();
  double nf = 0.0;
  for (int g = 0; g < R.G; ++g) nf += cloadd(R, D.pcpart + (size_t)g * PC_TERMS + 18);
  double px = 0.0, py = 0.0;
  for (int s = threadIdx.x; s < n; s += R.nt) {
    const double w = (cloadd(R, D.pcc + s) - mean) / nf;
    px += w * (double)(s % nw - r);
    py += w * (double)(s / nw - r);
  }
  double gx = block_sum(R, px);
  double gy = block_sum(R, py);
  if (threadIdx.x == 0) {
    gx *= 5.0;
    gy *= 5.0;
    if (fabs(gx) > 10.0) gx = 10.0 * gx / fabs(gx);
    if (fabs(gy) > 10.0) gy = 10.0 * gy / fabs(gy);
    double* Dk = D.a.pc_direction + 2 * (size_t)k;
    Dk[0] = 0.5 * gy + 0.9 * Dk[0];
    Dk[1] = 0.5 * gx + 0.9 * Dk[1];
  }
}

// ---------------------------------------------------------------------------
// e3PIE slice chain (e3PIE.py:89-173, updates :192-225).  W holds the
// (lambda, psm) fields of the slice being propagated; ESW keeps the exit
// wave of every slice of the current position.  Field f = il * npsm + ip.
// ---------------------------------------------------------------------------
template <int NP>
__device__ __forceinline__ void e3_rows_first(Run& R, int k, unsigned* slot, unsigned* zslot) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int nl = D.a.nlambda, npm = D.a.npsm, ns = D.a.nslice;
  const int No = D.a.No, nF = D.nF;
  const size_t No2 = (size_t)No * No, Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  const float2* __restrict__ O = (const float2*)D.a.object;
  const float2* __restrict__ P = (const float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  float2* __restrict__ E = D.ESW;
  const bool cb = D.a.propagator == PTY_FRAUNHOFER && ns == 1;
  clear_slot(R, zslot);
  float omx[MAX_SLICES];
#pragma unroll
  for (int s = 0; s < MAX_SLICES; ++s) omx[s] = 0.f;
  float pmax = 0.f;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    const size_t orow_off = (size_t)(row0 + r) * No + col0;
    float2 of[L], pv[L], q[L];
    load_row<NP>(R, of, O + (size_t)((f / npm) * ns + 0) * No2 + orow_off, R.t);
    load_row<NP>(R, pv, P + (size_t)(f * ns + 0) * Np2 + (size_t)r * NP, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) q[l] = make_float2(act ? cabs2(pv[l]) : 0.f, 0.f);
    lds_put<NP>(R.xa, islot, q, R.t);
    float2 v[L];
#pragma unroll
    for (int l = 0; l < L; ++l) v[l] = cmul(of[l], pv[l]);
    if (act) store_row<NP>(R, E + (size_t)f * Np2 + (size_t)r * NP, v, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (cb && ((r + l * LPF + R.t) & 1)) v[l] = make_float2(-v[l].x, -v[l].y);
    fft_fwd<NP>(v, R.tw, R.t, R.fx);
    if (act) store_row<NP>(R, W + f * Np2 + (size_t)r * NP, v, R.t);
    __syncthreads();
    if (ITEM_LEAD(act, frw)) {
      float sp[L];
#pragma unroll
      for (int l = 0; l < L; ++l) sp[l] = 0.f;
      for (int g = 0; g < nF; ++g) {
#pragma unroll
        for (int l = 0; l < L; ++l) sp[l] += R.xa[(lslot + g) * NP + l * LPF + R.t].x;
      }
#pragma unroll
      for (int l = 0; l < L; ++l) pmax = fmaxf(pmax, sp[l]);
#pragma unroll
      for (int s = 0; s < MAX_SLICES; ++s) {
        if (s < ns) {
#pragma unroll
          for (int l = 0; l < L; ++l) {
            float so = 0.f;
            for (int il = 0; il < nl; ++il)
              so += cabs2(cload(R, O + (size_t)(il * ns + s) * No2 + orow_off + l * LPF + R.t));
            omx[s] = fmaxf(omx[s], so);
          }
        }
      }
    }
    __syncthreads();
  ITEMS_END
  pmax = wave_max(pmax);
  if (R.lane == 0) atomic_fmax_slot(R, slot, 0, pmax);
#pragma unroll
  for (int s = 0; s < MAX_SLICES; ++s) {
    if (s < ns) {
      const float m = wave_max(omx[s]);
      if (R.lane == 0) atomic_fmax_slot(R, slot, 2 * s + 1, m);
    }
  }
}

// Forward chain, slice s >= 1: P_s = ifft2(fft2(esw_{s-1}) H) (the column
// pass already applied H / Np^2), stored into probe slice s; esw_s = P_s O_s;
// then the forward row transform of esw_s (with the detector checkerboard
// when s is the last slice).
template <int NP>
__device__ __forceinline__ void e3_rows_forward(Run& R, int k, int s, unsigned* slot) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int npm = D.a.npsm, ns = D.a.nslice;
  const int No = D.a.No, nF = D.nF;
  const size_t No2 = (size_t)No * No, Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  const float2* __restrict__ O = (const float2*)D.a.object;
  float2* __restrict__ P = (float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  float2* __restrict__ E = D.ESW;
  const bool cb = D.a.propagator == PTY_FRAUNHOFER && s == ns - 1;
  float pmax = 0.f;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    const size_t orow_off = (size_t)(row0 + r) * No + col0;
    float2 v[L], os[L], q[L];
    load_row<NP>(R, v, W + f * Np2 + (size_t)r * NP, R.t);
    load_row<NP>(R, os, O + (size_t)((f / npm) * ns + s) * No2 + orow_off, R.t);
    fft_inv<NP>(v, R.tw, R.t, R.fx);
    if (act) store_row<NP>(R, P + (size_t)(f * ns + s) * Np2 + (size_t)r * NP, v, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l) {
      q[l] = make_float2(act ? cabs2(v[l]) : 0.f, 0.f);
      v[l] = cmul(v[l], os[l]);
    }
    lds_put<NP>(R.xa, islot, q, R.t);
    if (act) store_row<NP>(R, E + (size_t)(s * nF + f) * Np2 + (size_t)r * NP, v, R.t);
#pragma unroll
    for (int l = 0; l < L; ++l)
      if (cb && ((r + l * LPF + R.t) & 1)) v[l] = make_float2(-v[l].x, -v[l].y);
    fft_fwd<NP>(v, R.tw, R.t, R.fx);
    if (act) store_row<NP>(R, W + f * Np2 + (size_t)r * NP, v, R.t);
    __syncthreads();
    if (ITEM_LEAD(act, frw)) {
      float sp[L];
#pragma unroll
      for (int l = 0; l < L; ++l) sp[l] = 0.f;
      for (int g = 0; g < nF; ++g) {
#pragma unroll
        for (int l = 0; l < L; ++l) sp[l] += R.xa[(lslot + g) * NP + l * LPF + R.t].x;
      }
#pragma unroll
      for (int l = 0; l < L; ++l) pmax = fmaxf(pmax, sp[l]);
    }
    __syncthreads();
  ITEMS_END
  pmax = wave_max(pmax);
  if (R.lane == 0) atomic_fmax_slot(R, slot, 2 * s, pmax);
}

// Update of slice su from the difference wave: `fromDetector` selects the
// detector back-propagation (Fraunhofer scale and checkerboard, difference to
// the last exit wave) or a slice back-propagation (difference to esw_su).
// Object slice su and probe slice su are rewritten; for su >= 1 the updated
// probe slice is forward-transformed into W for the next back-propagation.
template <int NP, bool MXL = false>
__device__ __forceinline__ void e3_rows_update(Run& R, int k, int su, const unsigned* slot, bool fromDetector) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF;
  const DevInst& D = *R.D;
  const int npm = D.a.npsm, ns = D.a.nslice;
  const int No = D.a.No, nF = D.nF;
  const size_t No2 = (size_t)No * No, Np2 = (size_t)NP * NP;
  const int row0 = D.a.positions[2 * k], col0 = D.a.positions[2 * k + 1];
  float2* __restrict__ O = (float2*)D.a.object;
  float2* __restrict__ P = (float2*)D.a.probe;
  float2* __restrict__ W = D.W;
  const float2* __restrict__ E = D.ESW;
  const bool cb = D.a.propagator == PTY_FRAUNHOFER;
  const float invN = 1.0f / (float)NP;
  float mP, mO;
  if (MXL) {
    // words 2 su, 2 su + 1, left in LDS by the column pass before this one
    mP = __uint_as_float(R.smx[3]);
    mO = __uint_as_float(R.smx[4]);
  } else {
    load_slot2(R, slot, 2 * su, 2 * su + 1, mP, mO);
  }
  const float invP = 1.0f / mP;
  const float invO = 1.0f / mO;
  const float bO = D.a.betaObject;
  const float bP = su >= 1 ? D.a.sliceBeta : D.a.betaProbe;
  const int esl = fromDetector ? ns - 1 : su;
  ITEMS_BEGIN(R, NP, r, f, act, islot, lslot, frw)
    (void)frw;
    const size_t orow_off = (size_t)(row0 + r) * No + col0;
    const size_t prow_off = (size_t)r * NP;
    float2 d[L], pv[L], ov[L], ev[L];
    load_row<NP>(R, d, W + f * Np2 + prow_off, R.t);
    load_row<NP>(R, pv, P + (size_t)(f * ns + su) * Np2 + prow_off, R.t);
    load_row<NP>(R, ev, E + (size_t)(esl * nF + f) * Np2 + prow_off, R.t);
    load_row<NP>(R, ov, O + (size_t)((f / npm) * ns + su) * No2 + orow_off, R.t);
    fft_inv<NP>(d, R.tw, R.t, R.fx);
    float2 tO[L], tP[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float2 u = d[l];
      if (fromDetector) {
        u = cscale(u, invN);
        if (cb && ((r + l * LPF + R.t) & 1)) u = make_float2(-u.x, -u.y);
      }
      const float2 dl = csub(u, ev[l]);
      tO[l] = cmul(make_float2(pv[l].x * invP, -pv[l].y * invP), dl);
      tP[l] = cmul(make_float2(ov[l].x * invO, -ov[l].y * invO), dl);
    }
    lds_put<NP>(R.xa, islot, tO, R.t);
    lds_put<NP>(R.xb, islot, tP, R.t);
    __syncthreads();
    if (act && f % npm == 0) {
      // object slice su of this lambda: sum over (lambda, psm) (axes (0, 2))
      float2 nv[L], acc[L];
#pragma unroll
      for (int l = 0; l < L; ++l) acc[l] = make_float2(0.f, 0.f);
      for (int g = 0; g < nF; ++g) {
#pragma unroll
        for (int l = 0; l < L; ++l) acc[l] = cadd(acc[l], R.xa[(lslot + g) * NP + l * LPF + R.t]);
      }
#pragma unroll
      for (int l = 0; l < L; ++l) nv[l] = cadd(ov[l], cscale(acc[l], bO));
      store_row<NP>(R, O + (size_t)((f / npm) * ns + su) * No2 + orow_off, nv, R.t);
    }
    // probe slice su of this field: sum over lambda (axes (0, 1))
    float2 nv[L];
#pragma unroll
    for (int l = 0; l < L; ++l) {
      float2 acc = make_float2(0.f, 0.f);
      for (int g = f % npm; g < nF; g += npm) acc = cadd(acc, R.xb[(lslot + g) * NP + l * LPF + R.t]);
      nv[l] = cadd(pv[l], cscale(acc, bP));
    }
    if (act) store_row<NP>(R, P + (size_t)(f * ns + su) * Np2 + prow_off, nv, R.t);
    if (su >= 1) {
      fft_fwd<NP>(nv, R.tw, R.t, R.fx);
      if (act) store_row<NP>(R, W + f * Np2 + prow_off, nv, R.t);
    }
    __syncthreads();
  ITEMS_END
}

// ---------------------------------------------------------------------------
// per-position drivers
// ---------------------------------------------------------------------------
template <int NP, bool ASP, int CWR, bool QN, bool QPH, bool PC, bool BAT>
__device__ __forceinline__ bool pie_position(Run& R, int k, int j, unsigned* slot, unsigned* zslot, double* ep,
                                             bool sync_after, int it, float* de) {
  const DevInst& D = *R.D;
  // only the angular-spectrum families take the field-loop forms: with them
  // the Fraunhofer family spilled more registers and lost 6 % on one-mode
  // instances, which have no field sums to save
  constexpr bool FLK = BAT && ASP && !QN && !PC;
  // the pass before the update reads the maxima into LDS: the constraint
  // pass of the per-size angular-spectrum kernels, the staged detector
  // column pass of the per-size Fraunhofer ones (in the batched kernel the
  // registers this holds through the pass cost more in spills than the
  // round trip it saves)
  constexpr bool MXF = !ASP && !BAT;
  constexpr bool MXU = ASP ? !BAT : MXF && col_staged<NP>();
  // per-size kernels whose waves own one (line, field) item per row pass
  // keep the item's patch and probe rows in registers from the first row
  // pass to the last (the batched kernel has no registers to spare)
  float2 ko[Line<NP>::L], kp[Line<NP>::L];
  const bool keep = !BAT && NP / R.G <= R.ipr / R.nFp;
  if (FLK && R.fl)
    pie_rows_first_fl<NP, ASP, QPH>(R, k, slot, zslot);
  else if (!BAT && keep)
    pie_rows_first<NP, QN, ASP, QPH, true, !BAT>(R, k, slot, zslot, ko, kp);
  else
    pie_rows_first<NP, QN, ASP, QPH, true>(R, k, slot, zslot, ko, kp);
  if (!cluster_sync(R)) return false;
  if constexpr (ASP) {
    tf_cols<NP, CWR, true>(R, false);
    if (!cluster_sync(R)) return false;
    asp_rows_constraint<NP, true, BAT ? 0 : 4>(R, k, j, ep, de, slot);
    if (!cluster_sync(R)) return false;
    tf_cols<NP, CWR, true>(R, true);
    if (!cluster_sync(R)) return false;
  } else {
    detector_cols_fraunhofer<NP, CWR, MXF && col_staged<NP>() ? (BAT ? 1 : 4) : 0>(R, k, j, ep, D.fs, de, slot);
    if (!cluster_sync(R)) return false;
  }
  if (FLK && R.fl)
    pie_rows_update_fl<NP, ASP, QPH>(R, k, slot);
  else
    if (!BAT && keep)
      pie_rows_update<NP, QN, ASP, QPH, true, !BAT, MXU>(R, k, slot, !sync_after, ko, kp);
    else
      pie_rows_update<NP, QN, ASP, QPH, true, false, MXU>(R, k, slot, !sync_after, ko, kp);
  if (sync_after && !cluster_sync(R)) return false;
  if constexpr (PC) {
    if (it >= D.a.pc_start) {
      if (D.a.pc_radius >= 2) {
        pcf_rows<NP>(R, k);
        if (!cluster_sync(R)) return false;
        pcf_cols<NP, CWR>(R);
        if (!cluster_sync(R)) return false;
        pcf_window_rows<NP>(R);
        if (!cluster_sync(R)) return false;
        pcf_direction_update(R, k);
      } else {
        pc_corr_pass<NP>(R, k);
        if (!cluster_sync(R)) return false;
        pc_direction_update(R, k);
      }
    }
  }
  if constexpr (QN) {
    qn_omax_pass<NP>(R, slot);
    if (!cluster_sync(R)) return false;
    qn_probe_rows<NP>(R, slot);
    if (!cluster_sync(R)) return false;
    if (D.a.engine == PTY_MQNEWTON) {
      if (!(D.a.flags & PTY_FLAG_MQN_MOMENTUM)) {
        adam_norm_pass(R);
        if (!cluster_sync(R)) return false;
      }
      adam_update_pass(R, it + 1, (unsigned long long)it * D.a.F + j);
      if (!cluster_sync(R)) return false;
    }
  }
  return true;
}

template <int NP, int CWR, bool BAT>
__device__ __forceinline__ bool e3pie_position(Run& R, int k, int j, unsigned* slot, unsigned* zslot, double* ep,
                                               float* de) {
  // per-size kernels up to Np = 128: every update pass takes its slice's
  // maxima from LDS, read by the column pass before it (final since the
  // forward chain); the Np = 256 kernel already spills, and the registers
  // held through the column pass cost it more than the round trip saved
  constexpr bool MX = !BAT && NP <= 128;
  const DevInst& D = *R.D;
  const int ns = D.a.nslice;
  e3_rows_first<NP>(R, k, slot, zslot);
  if (!cluster_sync(R)) return false;
  for (int s = 1; s < ns; ++s) {
    tf_cols<NP, CWR, false>(R, false);
    if (!cluster_sync(R)) return false;
    e3_rows_forward<NP>(R, k, s, slot);
    if (!cluster_sync(R)) return false;
  }
  detector_cols_fraunhofer<NP, CWR, MX ? 4 : 0>(R, k, j, ep, 1, de, slot + (size_t)2 * (ns - 1) * R.G);
  if (!cluster_sync(R)) return false;
  e3_rows_update<NP, MX>(R, k, ns - 1, slot, true);
  if (!cluster_sync(R)) return false;
  for (int s = ns - 1; s >= 1; --s) {
    tf_cols<NP, CWR, false, MX ? 4 : 0>(R, true, slot + (size_t)2 * (s - 1) * R.G);
    if (!cluster_sync(R)) return false;
    e3_rows_update<NP, MX>(R, k, s - 1, slot, false);
    if (!cluster_sync(R)) return false;
  }
  return true;
}

// Iterations [it0, it1); the first iteration starts at position jb and the
// last one stops before position je (full iterations: jb = 0, je >= F).  The
// error of an iteration is reduced when its last position has run.
template <int NP, int FAM, int CWR, bool BAT = false>
__device__ __forceinline__ void run_iterations(Run& R, int it0, int it1, int jb, int je) {
  constexpr int BF = base_family(FAM);  // pass family without the position correction
  constexpr bool PC = BF != FAM;
  const DevInst& D = *R.D;
  const int F = D.a.F;
  const int engine = D.a.engine;
  const bool momentum = (engine == PTY_MPIE || engine == PTY_MULTIPIE) && D.a.momentum_mask != nullptr;
  // instances of a batch may have fewer iterations than the launch covers
  if (it1 > D.a.n_iter) it1 = D.a.n_iter;
  unsigned q = 0;
  for (int it = it0; it < it1; ++it) {
    const int* order = D.a.order + (size_t)it * F;
    double* ep = D.errpart + (size_t)(it & 1) * F * R.G;
    const int j0 = it == it0 ? (jb < F ? jb : F) : 0;
    const int j1 = it == it1 - 1 ? (je < F ? je : F) : F;
    for (int j = j0; j < j1; ++j, ++q) {
      const int k = order[j];
      if constexpr (PC) {
        // positions are checked at plan creation; with the position
        // correction the caller rewrites them between launches, so they are
        // checked again here; every workgroup of the cluster reads the same
        // value and leaves at the same point
        const int pr = D.a.positions[2 * k], pc = D.a.positions[2 * k + 1];
        if (pr < 0 || pc < 0 || pr + NP > D.a.No || pc + NP > D.a.No) {
          if (threadIdx.x == 0) __hip_atomic_store(D.ctrl + CTRL_BADPOS, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          return;
        }
      }
      unsigned* slot = D.mx + (size_t)(q & 3u) * SLOT_WORDS * R.G;
      unsigned* zslot = D.mx + (size_t)((q + 2u) & 3u) * SLOT_WORDS * R.G;
      const bool mom = BF != FAM_E3PIE && momentum && D.a.momentum_mask[(size_t)it * F + j];
      // The barrier after the update pass orders this position's object
      // writes before the next position's patch reads by other workgroups;
      // every other cross-workgroup dependency between the two (probe rows,
      // W rows, maxima, error partials) stays inside one workgroup or is
      // ordered by an earlier barrier.  When the next patch in this launch
      // does not overlap this one and no momentum step follows, the PIE
      // engines skip it.
      bool sync_after = true;
      if constexpr (BF == FAM_PIE_FRAUNHOFER || BF == FAM_PIE_ASP || BF == FAM_PIE_SASP) {
        int kn = -1;
        if (j + 1 < j1)
          kn = order[j + 1];
        else if (j1 == F && it + 1 < it1)
          kn = D.a.order[(size_t)(it + 1) * F];
        if (!mom && kn >= 0 && !(PC && it >= D.a.pc_start)) {
          const int dr = D.a.positions[2 * k] - D.a.positions[2 * kn];
          const int dc = D.a.positions[2 * k + 1] - D.a.positions[2 * kn + 1];
          sync_after = !(dr >= NP || -dr >= NP || dc >= NP || -dc >= NP);
        }
      }
      // the detector error of the instance's last iteration, when asked for
      float* de = D.a.detector_error && it == D.a.n_iter - 1
                      ? D.a.detector_error + (size_t)k * NP * NP
                      : nullptr;
      bool ok;
      if constexpr (BF == FAM_E3PIE)
        ok = e3pie_position<NP, CWR, BAT>(R, k, j, slot, zslot, ep, de);
      else
        ok = pie_position<NP, BF == FAM_PIE_ASP || BF == FAM_QN_ASP || BF == FAM_PIE_SASP, CWR,
                          BF == FAM_QN_FRAUNHOFER || BF == FAM_QN_ASP, BF == FAM_PIE_SASP, PC, BAT>(
            R, k, j, slot, zslot, ep, sync_after, it, de);
      if (!ok) return;
      if (mom) {
        momentum_pass<NP>(R);
        if (!cluster_sync(R)) return;
      }
    }
    if (j1 == F && R.g == 0) {
      // getErrorMetrics (BaseEngine.py:819-841): errorAtPos = sum|I - Iest| /
      // (energyAtPos + 1e-20) in FP64, error = sum over positions
      double acc = 0.0;
      for (int jj = threadIdx.x; jj < F; jj += R.nt) {
        const int kk = order[jj];
        double s = 0.0;
        for (int g = 0; g < R.G; ++g) s += cloadd(R, ep + (size_t)jj * R.G + g);
        const float en = D.a.energy[kk] + 1e-20f;
        const double e = s / (double)en;
        D.a.err_at_pos[(size_t)it * F + kk] = e;
        acc += e;
      }
      const double tot = block_sum(R, acc);
      if (threadIdx.x == 0) D.a.err_iter[it] = tot;
    }
  }
}

// Counts the cluster's workgroups on each XCD for the XCD barrier: every
// workgroup adds itself to the count of its XCD, one counter barrier makes
// the counts final, then thread 0 keeps its XCD's count and the number of
// XCDs the cluster occupies.  Returns false when that barrier timed out.
__device__ __forceinline__ bool xcd_barrier_setup(Run& R) {
  unsigned* ctrl = R.D->ctrl;
  if (threadIdx.x == 0) {
    R.xcd = xcc_id();
    __hip_atomic_fetch_add(ctrl + CTRL_XPOP + R.xcd, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (!cluster_sync(R)) return false;
  if (threadIdx.x == 0) {
    unsigned nx = 0u;
    for (int x = 0; x < N_XCD; ++x) {
      const unsigned c = __hip_atomic_load(ctrl + CTRL_XPOP + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      nx += c > 0u ? 1u : 0u;
      if ((unsigned)x == R.xcd) R.xpop = c;
    }
    R.nx = nx;
  }
  R.xbar = true;
  return true;
}

// Per-workgroup set-up for probe size NP on a workgroup of NT threads:
// lane roles, LDS views, the twiddles of length NP.
template <int NP, int NT, int CWR>
__device__ __forceinline__ void init_run(Run& R, const DevInst* D, int g, float2* s_tw, double* s_red, float2* s_x,
                                         float2* s_fx, int* s_fail, unsigned* s_mx) {
  R.D = D;
  R.g = g;
  R.G = D->G;
  R.nF = D->nF;
  R.nFp = D->nFp;
  R.wave = threadIdx.x >> 6;
  R.lane = threadIdx.x & 63;
  R.t = R.lane & (Line<NP>::LPF - 1);
  R.sub = R.lane / Line<NP>::LPF;
  R.gen = 0u;
  R.tw = s_tw;
  R.red = s_red;
  R.nt = NT;
  R.nwave = NT / 64;
  R.ipr = (NT / 64) * Line<NP>::SUB;
  R.xa = s_x;
  R.xb = s_x + R.ipr * NP;
  R.fx = s_fx + R.wave * wave_scratch<NP, CWR>();
  R.fail = s_fail;
  R.smx = s_mx;
  if (threadIdx.x < 3) s_mx[threadIdx.x] = 0u;
  R.loc = false;
  R.fl = false;
  R.xbar = false;
  R.xgen = 0u;
  R.xcd = R.xpop = R.nx = 0u;
  R.lcix = nullptr;
  for (int i = threadIdx.x; i < NP / 2; i += NT) s_tw[i] = g_twiddle[Line<NP>::LOG2 - 4][i];
  __syncthreads();
}

// The packed-row column codes of the workgroup's instance in LDS (when they
// fit): the row passes look them up before every exchange-row access, and
// from LDS that lookup does not add a memory round trip to the row's load.
// families whose exchange rows can be packed (the angular-spectrum PIE ones)
constexpr bool fam_packed(int fam) {
  return fam == FAM_PIE_ASP || fam == FAM_PIE_SASP || fam == FAM_PC_ASP || fam == FAM_PC_SASP;
}
template <int NP, int NT, int CAP = PACK_LDS>
__device__ __forceinline__ void cix_to_lds(Run& R, short* s_cix) {
  const DevInst& D = *R.D;
  const int n = D.cix ? (D.nF / D.fpl) * NP : 0;
  if (n > 0 && n <= CAP) {
    for (int i = threadIdx.x; i < n; i += NT) s_cix[i] = D.cix[i];
    __syncthreads();
    R.lcix = s_cix;
  }
}

// One kernel per probe size and pass family (PIE with the Fraunhofer
// detector, PIE with ASP, e3PIE): each instantiation carries only its own
// passes and gets its own register allocation and workgroup size.
template <int NP, int FAM>
__global__ __launch_bounds__(Cfg<NP>::NT) void pty_iterate_kernel(const DevInst* __restrict__ insts,
                                                                  const int2* __restrict__ bmap, int it0, int it1,
                                                                  int jb, int je, unsigned* claim, int nslots) {
  __shared__ float2 s_tw[256];
  __shared__ double s_red[Cfg<NP>::NW];
  __shared__ float2 s_x[Cfg<NP>::XBUF];
  __shared__ float2 s_fx[Cfg<NP>::NW * wave_scratch<NP, 0>()];
  __shared__ int s_fail;
  __shared__ unsigned s_mx[5];
  __shared__ short s_cix[PACK_LDS];
  const int2 bm = bmap[blockIdx.x];
  if (bm.x < 0) return;
  Run R;
  init_run<NP, Cfg<NP>::NT, 0>(R, insts + bm.x, bm.y, s_tw, s_red, s_x, s_fx, &s_fail, s_mx);
  cix_to_lds<NP, Cfg<NP>::NT>(R, s_cix);
  if (R.G > 1 && !xcd_barrier_setup(R)) return;
  run_iterations<NP, FAM, 0>(R, it0, it1, jb, je);
}

// Batched launch: every instance of one pass family with Np <= 256 runs in
// one persistent launch of BATCH_NT-thread workgroups, whatever its probe
// size; each workgroup branches once on its instance's Np.  One launch of
// uniform workgroups is what makes the co-residency of every cluster a
// property of the launch size alone (nblocks <= co-resident capacity).
#define BATCH_NT 512
#define BATCH_MAX_NP 256
#ifndef BATCH_CW
#define BATCH_CW 1  // column width of the batched kernel (fewest registers)
#endif

template <int NP>
struct BatchCfg {
  static constexpr int NW = BATCH_NT / 64;
  static constexpr int IPR = NW * Line<NP>::SUB;
  static constexpr int XROW = 2 * IPR * NP;
  static constexpr int XCOL = NW * NP * Col<NP, BATCH_CW>::CW / 2;
  static constexpr int XSTG = NW * Col<NP, BATCH_CW>::CW * (2 * NP + 32);  // staged column rounds
  static constexpr int XRC = XROW > XCOL ? XROW : XCOL;
  static constexpr int XBUF = XRC > XSTG ? XRC : XSTG;
};
constexpr int cmax(int a, int b) { return a > b ? a : b; }
constexpr int BATCH_XBUF =
    cmax(cmax(cmax(BatchCfg<16>::XBUF, BatchCfg<32>::XBUF), cmax(BatchCfg<64>::XBUF, BatchCfg<128>::XBUF)),
         BatchCfg<256>::XBUF);
constexpr int BATCH_FX = cmax(cmax(cmax(wave_scratch<16, BATCH_CW>(), wave_scratch<32, BATCH_CW>()),
                                   cmax(wave_scratch<64, BATCH_CW>(), wave_scratch<128, BATCH_CW>())),
                              wave_scratch<256, BATCH_CW>());

template <int FAM>
__global__ __launch_bounds__(BATCH_NT) __attribute__((amdgpu_waves_per_eu(4))) void pty_batch_kernel(
    const DevInst* __restrict__ insts, const int2* __restrict__ bmap, int it0, int it1, int jb, int je,
    unsigned* claim, int nslots) {
  __shared__ float2 s_tw[256];
  __shared__ double s_red[BATCH_NT / 64];
  __shared__ float2 s_x[BATCH_XBUF];
  __shared__ float2 s_fx[(BATCH_NT / 64) * BATCH_FX];
  __shared__ int s_fail;
  __shared__ unsigned s_mx[5];
  constexpr int CIX = fam_packed(FAM) ? PACK_LDS : 1;
  __shared__ short s_cix[CIX];
  __shared__ int s_blk;
  // The plan packs every cluster onto one XCD: bmap entry N_XCD * s + x is
  // slot s of XCD x.  A workgroup reads the XCD it runs on (XCC_ID hardware
  // register) and claims the next slot of that XCD, so the placement holds
  // whatever XCD the dispatcher's round robin starts from.
  if (threadIdx.x == 0) {
    const unsigned xcc = (unsigned)__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 7u;
    const unsigned slot = __hip_atomic_fetch_add(claim + 32 * xcc, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_blk = slot < (unsigned)nslots ? (int)(N_XCD * slot + xcc) : -1;
  }
  __syncthreads();
  const int blk = s_blk;
  if (blk < 0) {
    // more workgroups reached this XCD than it has slots, so some cluster
    // misses a member: stop every cluster of the launch
    for (int b = threadIdx.x; b < N_XCD * nslots; b += BATCH_NT) {
      const int2 e = bmap[b];
      if (e.x >= 0 && e.y == 0) {
        __hip_atomic_store(insts[e.x].ctrl + CTRL_PLACE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(insts[e.x].ctrl + CTRL_FAIL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    return;
  }
  const int2 bm = bmap[blk];
  if (bm.x < 0) return;
  const DevInst* D = insts + bm.x;
  Run R;
  switch (D->a.Np) {
    case 16:
      init_run<16, BATCH_NT, BATCH_CW>(R, D, bm.y, s_tw, s_red, s_x, s_fx, &s_fail, s_mx);
      cix_to_lds<16, BATCH_NT, CIX>(R, s_cix);
      R.loc = true;
      R.fl = D->fl && 16 / R.G >= R.ipr;
      run_iterations<16, FAM, BATCH_CW, true>(R, it0, it1, jb, je);
      break;
    case 32:
      init_run<32, BATCH_NT, BATCH_CW>(R, D, bm.y, s_tw, s_red, s_x, s_fx, &s_fail, s_mx);
      cix_to_lds<32, BATCH_NT, CIX>(R, s_cix);
      R.loc = true;
      R.fl = D->fl && 32 / R.G >= R.ipr;
      run_iterations<32, FAM, BATCH_CW, true>(R, it0, it1, jb, je);
      break;
    case 64:
      init_run<64, BATCH_NT, BATCH_CW>(R, D, bm.y, s_tw, s_red, s_x, s_fx, &s_fail, s_mx);
      cix_to_lds<64, BATCH_NT, CIX>(R, s_cix);
      R.loc = true;
      R.fl = D->fl && 64 / R.G >= R.ipr;
      run_iterations<64, FAM, BATCH_CW, true>(R, it0, it1, jb, je);
      break;
    case 128:
      init_run<128, BATCH_NT, BATCH_CW>(R, D, bm.y, s_tw, s_red, s_x, s_fx, &s_fail, s_mx);
      cix_to_lds<128, BATCH_NT, CIX>(R, s_cix);
      R.loc = true;
      R.fl = D->fl && 128 / R.G >= R.ipr;
      run_iterations<128, FAM, BATCH_CW, true>(R, it0, it1, jb, je);
      break;
    case 256:
      init_run<256, BATCH_NT, BATCH_CW>(R, D, bm.y, s_tw, s_red, s_x, s_fx, &s_fail, s_mx);
      cix_to_lds<256, BATCH_NT, CIX>(R, s_cix);
      R.loc = true;
      R.fl = D->fl && 256 / R.G >= R.ipr;
      run_iterations<256, FAM, BATCH_CW, true>(R, it0, it1, jb, je);
      break;
  }
}

// clears the control block of every instance (one block each) and the
// XCD slot counters of the batched launches
__global__ void pty_reset_ctrl_kernel(const DevInst* __restrict__ insts, unsigned* claim, int nclaim) {
  unsigned* c = insts[blockIdx.x].ctrl;
  for (int i = threadIdx.x; i < CTRL_WORDS; i += blockDim.x) c[i] = 0u;
  unsigned* m = insts[blockIdx.x].mx;
  const int nm = 4 * SLOT_WORDS * insts[blockIdx.x].G;
  for (int i = threadIdx.x; i < nm; i += blockDim.x) m[i] = 0u;
  unsigned* b = insts[blockIdx.x].bar;
  for (int i = threadIdx.x; i < insts[blockIdx.x].G; i += blockDim.x) b[i] = 0u;
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < nclaim; i += blockDim.x) claim[i] = 0u;
}

// ---------------------------------------------------------------------------
// layout preparation kernels (plan creation only)
// ---------------------------------------------------------------------------
// dst[f][n][p] = src[f][brev(n)][brev(p)]: the measured frame in the order
// the column pass holds the transformed field (rows and columns bit-reversed)
__global__ void pty_permute_frames_kernel(const float* __restrict__ src, float* __restrict__ dst, int F, int np,
                                          int lg) {
  const size_t np2 = (size_t)np * np;
  const size_t total = (size_t)F * np2;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const size_t f = i / np2;
    const int rem = (int)(i % np2);
    const int n = rem / np, p = rem % np;
    dst[i] = src[f * np2 + (size_t)brev(n, lg) * np + brev(p, lg)];
  }
}

// ht[n][p] = tf[brev(n)][brev(p)] / np^2 (np^2 is a power of two: exact)
// live[t][p] = 1 when column p of transfer function t (kernel layout) has a
// nonzero entry
__global__ void pty_live_columns_kernel(const float2* __restrict__ ht, unsigned char* __restrict__ live, int np,
                                        int ntf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ntf * np) return;
  const int t = i / np, p = i - t * np;
  const float2* h = ht + (size_t)t * np * np;
  unsigned char any = 0;
  for (int r = 0; r < np; ++r) {
    const float2 v = h[(size_t)r * np + p];
    any |= (v.x != 0.f || v.y != 0.f);
  }
  live[i] = any;
}

// lgrp[t][m] = the m-th column group (cw adjacent columns) of transfer function
// t holding a live column, nlive[1 + t] their number, nlive[0] the largest
// such number; with cix, the packed-row code of every column (load_row_live):
// m * cw + c for column c of the m-th live group when that column is live,
// -1 - (m * cw + c) when it is dead, PACK_DEAD for the columns of dead
// groups; one thread per transfer function
__global__ void pty_live_groups_kernel(const unsigned char* __restrict__ live, short* __restrict__ lgrp,
                                       int* __restrict__ nlive, short* __restrict__ cix, int np, int cw, int ntf) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int ng = np / cw;
  int most = 0;
  for (int t = 0; t < ntf; ++t) {
    int n = 0;
    for (int gi = 0; gi < ng; ++gi) {
      int any = 0;
      for (int c = 0; c < cw; ++c) any |= live[(size_t)t * np + gi * cw + c];
      if (cix) {
        for (int c = 0; c < cw; ++c) {
          const int pc = gi * cw + c, j = n * cw + c;
          cix[(size_t)t * np + pc] = (short)(!any ? PACK_DEAD : live[(size_t)t * np + pc] ? j : -1 - j);
        }
      }
      if (any) lgrp[(size_t)t * ng + n++] = (short)gi;
    }
    nlive[1 + t] = n;
    most = n > most ? n : most;
  }
  nlive[0] = most;
}

__global__ void pty_permute_tf_kernel(const float2* __restrict__ tf, float2* __restrict__ ht, int np, int lg) {
  const int total = np * np;
  const float sc = 1.0f / ((float)np * (float)np);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
    const int n = i / np, p = i % np;
    ht[i] = cscale(tf[brev(n, lg) * np + brev(p, lg)], sc);
  }
}

// ---------------------------------------------------------------------------
// standalone propagators and FFT (Operators.object2detector/detector2object)
// ---------------------------------------------------------------------------
// mode bits for pty_lines_kernel
#define LN_COLS 1        // process columns instead of rows
#define LN_INVERSE 2     // conjugate kernel
#define LN_PRE_CB 4      // multiply by (-1)^(y+x) before the transform
#define LN_POST_CB 8     // multiply by (-1)^(y+x) after the transform
#define LN_SCALE 16      // multiply by 1/Np after the transform

// Forward-structured (DIF) transform of every line of every field, written
// back in natural order (bit-reversal applied on the store).
template <int NP>
__global__ __launch_bounds__(NTHREADS) void pty_lines_kernel(float2* __restrict__ data, int nfields, int mode) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF, SUB = Line<NP>::SUB, LG = Line<NP>::LOG2;
  __shared__ float2 s_tw[256];
  __shared__ float2 s_fx[NWAVE * fft_scratch<NP, 1>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane & (LPF - 1), sub = lane / LPF;
  const bool inv = mode & LN_INVERSE;
  for (int i = threadIdx.x; i < NP / 2; i += NTHREADS) {
    const float2 w = g_twiddle[LG - 4][i];
    s_tw[i] = inv ? conjf2(w) : w;
  }
  __syncthreads();
  const int line = (blockIdx.x * NWAVE + wave) * SUB + sub;
  if (line >= nfields * NP) return;
  const int f = line / NP, li = line % NP;
  float2* base = data + (size_t)f * NP * NP;
  const bool cols = mode & LN_COLS;
  float2 v[L];
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int n = l * LPF + t;
    const size_t off = cols ? (size_t)n * NP + li : (size_t)li * NP + n;
    float2 x = base[off];
    if ((mode & LN_PRE_CB) && ((n + li) & 1)) x = make_float2(-x.x, -x.y);
    v[l] = x;
  }
  fft_fwd<NP>(v, s_tw, t, s_fx + wave * fft_scratch<NP, 1>());
  const float sc = 1.0f / (float)NP;
#pragma unroll
  for (int l = 0; l < L; ++l) {
    const int n = brev(l * LPF + t, LG);
    float2 x = v[l];
    if (mode & LN_SCALE) x = cscale(x, sc);
    if ((mode & LN_POST_CB) && ((n + li) & 1)) x = make_float2(-x.x, -x.y);
    const size_t off = cols ? (size_t)n * NP + li : (size_t)li * NP + n;
    base[off] = x;
  }
}

// ASP column pass for the standalone propagator: natural-order rows were
// forward transformed in place (bit-reversed x); transform the columns,
// multiply by tf (natural FFT order, read through the bit reversal), inverse
// column transform.
template <int NP>
__global__ __launch_bounds__(NTHREADS) void pty_asp_cols_kernel(float2* __restrict__ data, int nfields,
                                                                const float2* __restrict__ tf, int conj) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF, SUB = Line<NP>::SUB, LG = Line<NP>::LOG2;
  __shared__ float2 s_tw[256];
  __shared__ float2 s_fx[NWAVE * fft_scratch<NP, 1>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane & (LPF - 1), sub = lane / LPF;
  for (int i = threadIdx.x; i < NP / 2; i += NTHREADS) s_tw[i] = g_twiddle[LG - 4][i];
  __syncthreads();
  const int line = (blockIdx.x * NWAVE + wave) * SUB + sub;
  if (line >= nfields * NP) return;
  const int f = line / NP, p = line % NP;
  float2* base = data + (size_t)f * NP * NP;
  float2 v[L];
  load_col<NP>(v, base, p, t);
  fft_fwd<NP>(v, s_tw, t, s_fx + wave * fft_scratch<NP, 1>());
  const float sc = 1.0f / ((float)NP * (float)NP);
#pragma unroll
  for (int l = 0; l < L; ++l) {
    float2 h = tf[(size_t)brev(l * LPF + t, LG) * NP + brev(p, LG)];
    if (conj) h = conjf2(h);
    v[l] = cmul(v[l], cscale(h, sc));
  }
  fft_inv<NP>(v, s_tw, t, s_fx + wave * fft_scratch<NP, 1>());
  store_col<NP>(base, v, p, t);
}

template <int NP>
__global__ __launch_bounds__(NTHREADS) void pty_rows_fwd_kernel(float2* __restrict__ data, int nfields) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF, SUB = Line<NP>::SUB;
  __shared__ float2 s_tw[256];
  __shared__ float2 s_fx[NWAVE * fft_scratch<NP, 1>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane & (LPF - 1), sub = lane / LPF;
  for (int i = threadIdx.x; i < NP / 2; i += NTHREADS) s_tw[i] = g_twiddle[Line<NP>::LOG2 - 4][i];
  __syncthreads();
  const int line = (blockIdx.x * NWAVE + wave) * SUB + sub;
  if (line >= nfields * NP) return;
  float2* row = data + (size_t)line * NP;
  float2 v[L];
  load_row<NP>(v, row, t);
  fft_fwd<NP>(v, s_tw, t, s_fx + wave * fft_scratch<NP, 1>());
  store_row<NP>(row, v, t);
}

template <int NP>
__global__ __launch_bounds__(NTHREADS) void pty_rows_inv_kernel(float2* __restrict__ data, int nfields) {
  constexpr int L = Line<NP>::L, LPF = Line<NP>::LPF, SUB = Line<NP>::SUB;
  __shared__ float2 s_tw[256];
  __shared__ float2 s_fx[NWAVE * fft_scratch<NP, 1>()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = lane & (LPF - 1), sub = lane / LPF;
  for (int i = threadIdx.x; i < NP / 2; i += NTHREADS) s_tw[i] = g_twiddle[Line<NP>::LOG2 - 4][i];
  __syncthreads();
  const int line = (blockIdx.x * NWAVE + wave) * SUB + sub;
  if (line >= nfields * NP) return;
  float2* row = data + (size_t)line * NP;
  float2 v[L];
  load_row<NP>(v, row, t);
  fft_inv<NP>(v, s_tw, t, s_fx + wave * fft_scratch<NP, 1>());
  store_row<NP>(row, v, t);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Instances with the same probe size and pass family form a launch class:
// one kernel instantiation, one block map.  In a plan of several instances
// (or for an instance flagged PTY_FLAG_BATCHED) every instance with
// Np <= BATCH_MAX_NP joins the batched class of its family (np = 0), whose
// kernel serves every probe size.  Classes run one after another on the
// caller's stream, so each launch alone must fit the device.
struct LaunchClass {
  int np = 0;  // probe size; 0 for the batched class
  int fam = 0;
  int nblocks = 0;
  int nthreads = 0;
  bool cooperative = false;  // some instance of the class has G > 1
  int2* d_bmap = nullptr;
  unsigned* d_claim = nullptr;  // batched class: N_XCD slot counters, 32 words apart
  int nslots = 0;               // batched class: slots per XCD (nblocks / N_XCD)
  hipGraphNode_t node = nullptr;
};

struct pty_plan {
  int device = 0;
  int ninst = 0;
  int nblocks = 0;
  int max_iter = 0;
  DevInst* d_insts = nullptr;
  unsigned* d_claims = nullptr;  // slot counters of every class, cleared with the control blocks
  int nclaim = 0;
  std::vector<DevInst> h_insts;
  std::vector<LaunchClass> classes;
  std::vector<void*> allocs;
  // per-iteration graph: reset node, then one kernel node per class
  bool use_graph = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

template <int NP>
static const void* iterate_kernel_fam(int fam) {
  switch (fam) {
    case FAM_PIE_FRAUNHOFER: return (const void*)pty_iterate_kernel<NP, FAM_PIE_FRAUNHOFER>;
    case FAM_PIE_ASP: return (const void*)pty_iterate_kernel<NP, FAM_PIE_ASP>;
    case FAM_PIE_SASP: return (const void*)pty_iterate_kernel<NP, FAM_PIE_SASP>;
    case FAM_PC_FRAUNHOFER: return (const void*)pty_iterate_kernel<NP, FAM_PC_FRAUNHOFER>;
    case FAM_PC_ASP: return (const void*)pty_iterate_kernel<NP, FAM_PC_ASP>;
    case FAM_PC_SASP: return (const void*)pty_iterate_kernel<NP, FAM_PC_SASP>;
    case FAM_E3PIE: return (const void*)pty_iterate_kernel<NP, FAM_E3PIE>;
    case FAM_QN_FRAUNHOFER: return (const void*)pty_iterate_kernel<NP, FAM_QN_FRAUNHOFER>;
    case FAM_QN_ASP: return (const void*)pty_iterate_kernel<NP, FAM_QN_ASP>;
  }
  return nullptr;
}

static const void* iterate_kernel_ptr(int np, int fam) {
  switch (np) {
    case 0:
      switch (fam) {
        case FAM_PIE_FRAUNHOFER: return (const void*)pty_batch_kernel<FAM_PIE_FRAUNHOFER>;
        case FAM_PIE_ASP: return (const void*)pty_batch_kernel<FAM_PIE_ASP>;
        case FAM_PIE_SASP: return (const void*)pty_batch_kernel<FAM_PIE_SASP>;
        case FAM_E3PIE: return (const void*)pty_batch_kernel<FAM_E3PIE>;
        case FAM_QN_FRAUNHOFER: return (const void*)pty_batch_kernel<FAM_QN_FRAUNHOFER>;
        case FAM_QN_ASP: return (const void*)pty_batch_kernel<FAM_QN_ASP>;
      }
      return nullptr;
    case 16: return iterate_kernel_fam<16>(fam);
    case 32: return iterate_kernel_fam<32>(fam);
    case 64: return iterate_kernel_fam<64>(fam);
    case 128: return iterate_kernel_fam<128>(fam);
    case 256: return iterate_kernel_fam<256>(fam);
    case 512: return iterate_kernel_fam<512>(fam);
  }
  return nullptr;
}

static int family_of_base(const pty_instance& a);
static int family_of(const pty_instance& a) {
  const int f = family_of_base(a);
  if (!a.pc_direction) return f;
  return f == FAM_PIE_FRAUNHOFER ? FAM_PC_FRAUNHOFER : f == FAM_PIE_ASP ? FAM_PC_ASP : f == FAM_PIE_SASP ? FAM_PC_SASP : f;
}
static int family_of_base(const pty_instance& a) {
  if (a.engine == PTY_E3PIE) return FAM_E3PIE;
  if (a.engine == PTY_QNEWTON || a.engine == PTY_MQNEWTON) return prop_asp(a.propagator) ? FAM_QN_ASP : FAM_QN_FRAUNHOFER;
  if (a.propagator == PTY_SCALED_ASP || a.propagator == PTY_SCALED_POLYCHROME_ASP) return FAM_PIE_SASP;
  return prop_asp(a.propagator) ? FAM_PIE_ASP : FAM_PIE_FRAUNHOFER;
}

namespace {

bool g_tw_ready[64];

int ensure_twiddles() {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(PTY_EINVAL, "device index %d out of range", dev);
  if (g_tw_ready[dev]) return PTY_OK;
  static float2 host[6][256];
  for (int lg = 4; lg <= 9; ++lg) {
    const int n = 1 << lg;
    for (int j = 0; j < n / 2; ++j) {
      const double a = -2.0 * 3.14159265358979323846 * (double)j / (double)n;
      host[lg - 4][j] = make_float2((float)cos(a), (float)sin(a));
    }
  }
  HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(g_twiddle), host, sizeof host, 0, hipMemcpyHostToDevice));
  g_tw_ready[dev] = true;
  return PTY_OK;
}

int dalloc(pty_plan* p, void** out, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(out, bytes);
  if (e != hipSuccess) return set_err(PTY_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  p->allocs.push_back(*out);
  return PTY_OK;
}

void free_plan(pty_plan* p) {
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  p->exec = nullptr;
  p->graph = nullptr;
  p->classes.clear();
  for (void* a : p->allocs) (void)hipFree(a);
  p->allocs.clear();
}

// whether an instance runs in the batched class of its family
// Co-resident workgroups of the batched kernel of family fam on one XCD
// (gfx950: N_XCD XCDs of equal CU count).
int xcd_capacity(int fam, int* cap) {
  int dev = 0, cus = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, iterate_kernel_ptr(0, fam), BATCH_NT, 0));
  *cap = per_cu * cus / N_XCD;
  return PTY_OK;
}

bool joins_batch(const pty_instance& a, int ninst) {
  // the position correction runs in its own per-size kernels only
  return a.Np <= BATCH_MAX_NP && !a.pc_direction && (ninst > 1 || (a.flags & PTY_FLAG_BATCHED));
}

int auto_workgroups(const pty_instance& a, int ninst, int nt) {
  if (ninst > 1) return 1;
  int nF = a.engine == PTY_E3PIE ? a.nlambda * a.npsm : a.nlambda * a.nosm * a.npsm * a.nslice;
  int nFp = 1;
  while (nFp < nF) nFp *= 2;
  const int lines_per_round = slots_rt(a.Np, nt) / nFp;
  int g = a.Np / (lines_per_round > 0 ? lines_per_round : 1);
  // e3PIE's slice chain runs 4 nslice - 1 passes per position; from
  // Np = 128 on, twice the workgroups (half the waves of each busy) shorten
  // them (measured at Np = 128, four slices: 81.8 -> 76.3 us per position)
  if (a.engine == PTY_E3PIE && a.Np >= 128 && 2 * g <= a.Np / 4) g *= 2;
  if (g < 1) g = 1;
  if (g > 128) g = 128;
  return g;
}

int validate(const pty_instance& a, int idx, int nt) {
  if (a.flags & ~(PTY_FLAG_BATCHED | PTY_FLAG_MQN_NADAM | PTY_FLAG_MQN_MOMENTUM))
    return set_err(PTY_EINVAL, "instance %d: unknown flags 0x%x", idx, a.flags);
  if ((a.flags & (PTY_FLAG_MQN_NADAM | PTY_FLAG_MQN_MOMENTUM)) && a.engine != PTY_MQNEWTON)
    return set_err(PTY_EINVAL, "instance %d: PTY_FLAG_MQN_* needs engine PTY_MQNEWTON", idx);
  if ((a.flags & PTY_FLAG_MQN_NADAM) && (a.flags & PTY_FLAG_MQN_MOMENTUM))
    return set_err(PTY_EINVAL, "instance %d: PTY_FLAG_MQN_NADAM and PTY_FLAG_MQN_MOMENTUM exclude each other", idx);
  if (a.engine < PTY_EPIE || a.engine > PTY_MQNEWTON) return set_err(PTY_EINVAL, "instance %d: unknown engine %d", idx, a.engine);
  if (a.pc_direction && (a.engine == PTY_E3PIE || a.engine == PTY_QNEWTON || a.engine == PTY_MQNEWTON))
    return set_err(PTY_EINVAL, "instance %d: position correction runs with the ePIE, mPIE and multiPIE engines", idx);
  if (a.pc_direction && a.pc_radius >= 2 && 2 * a.pc_radius + 1 > a.Np)
    return set_err(PTY_EINVAL, "instance %d: pc_radius %d needs 2 pc_radius + 1 <= Np", idx, a.pc_radius);
  if (a.pc_direction && (a.nlambda != 1 || a.nosm != 1 || a.nslice != 1))
    return set_err(PTY_EINVAL, "instance %d: position correction needs nlambda = nosm = nslice = 1", idx);
  if (a.engine == PTY_MQNEWTON && (!a.objectBuffer || !a.objectMomentum || !a.probeBuffer || !a.probeMomentum))
    return set_err(PTY_EINVAL, "instance %d: mqNewton needs buffer and momentum arrays", idx);
  if (a.propagator < PTY_FRAUNHOFER || a.propagator > PTY_SCALED_POLYCHROME_ASP)
    return set_err(PTY_EINVAL, "instance %d: unknown propagator %d", idx, a.propagator);
  if (log2_np(a.Np) < 0)
    return set_err(PTY_EINVAL, "instance %d: Np=%d unsupported (powers of two 16..512)", idx, a.Np);
  if (a.nlambda < 1 || a.nosm < 1 || a.npsm < 1 || a.nslice < 1)
    return set_err(PTY_EINVAL, "instance %d: mode counts must be >= 1", idx);
  if (a.No < a.Np) return set_err(PTY_EINVAL, "instance %d: No=%d < Np=%d", idx, a.No, a.Np);
  if (a.F < 1 || a.n_iter < 1) return set_err(PTY_EINVAL, "instance %d: F and n_iter must be >= 1", idx);
  if (!a.object || !a.probe || !a.ptychogram || !a.positions || !a.order || !a.energy || !a.err_at_pos || !a.err_iter)
    return set_err(PTY_EINVAL, "instance %d: missing array pointer", idx);
  if (a.propagator == PTY_ASP && a.nlambda > 1)
    return set_err(PTY_EINVAL, "instance %d: For multi-wavelength, set polychromeASP instead of ASP", idx);
  if (a.propagator == PTY_ASP && !a.tf) return set_err(PTY_EINVAL, "instance %d: ASP needs a transfer function", idx);
  if (a.propagator == PTY_POLYCHROME_ASP && !a.tf)
    return set_err(PTY_EINVAL, "instance %d: polychromeASP needs one transfer function per wavelength", idx);
  if (a.propagator == PTY_FRESNEL && !a.tf)
    return set_err(PTY_EINVAL, "instance %d: Fresnel needs the quadratic phase", idx);
  if (a.propagator == PTY_SCALED_ASP || a.propagator == PTY_SCALED_POLYCHROME_ASP) {
    if (!a.tf) return set_err(PTY_EINVAL, "instance %d: scaledASP needs Q2 and the checkerboarded Q1", idx);
    if (a.propagator == PTY_SCALED_ASP && a.nlambda > 1)
      return set_err(PTY_EINVAL, "instance %d: For multi-wavelength, scaledPolychromeASP is needed instead of scaledASP",
                     idx);
    if (a.engine == PTY_QNEWTON || a.engine == PTY_MQNEWTON || a.engine == PTY_E3PIE)
      return set_err(PTY_EINVAL, "instance %d: scaledASP runs with the ePIE, mPIE and multiPIE engines", idx);
  }
  const int slots = slots_rt(a.Np, nt);
  {
    // a column pass gives every field of a column group its own wave in one
    // round (the mode sums go through LDS), so the padded field count may
    // not exceed the waves of a workgroup
    const int nF = a.engine == PTY_E3PIE ? a.nlambda * a.npsm : a.nlambda * a.nosm * a.npsm * a.nslice;
    int nFp = 1;
    while (nFp < nF) nFp *= 2;
    if (nFp > nt / 64)
      return set_err(PTY_EINVAL, "instance %d: %d exit-wave fields exceed the %d waves of a workgroup for Np=%d", idx, nF,
                     nt / 64, a.Np);
  }
  if (a.No % 2 != 0) return set_err(PTY_EINVAL, "instance %d: No=%d must be even", idx, a.No);
  if (a.engine == PTY_E3PIE) {
    if (a.nosm != 1) return set_err(PTY_EINVAL, "instance %d: e3PIE needs nosm == 1", idx);
    if (a.nslice > MAX_SLICES) return set_err(PTY_EINVAL, "instance %d: nslice > %d", idx, MAX_SLICES);
    if (a.propagator != PTY_FRAUNHOFER && a.propagator != PTY_FRAUNHOFER_SHIFTED)
      return set_err(PTY_EINVAL, "instance %d: e3PIE supports the Fraunhofer detector propagator only", idx);
    if (!a.tf) return set_err(PTY_EINVAL, "instance %d: e3PIE needs the slice transfer function", idx);
    if (a.nlambda * a.npsm > slots)
      return set_err(PTY_EINVAL, "instance %d: more than %d modes for Np=%d", idx, slots, a.Np);
  } else {
    const int nF = a.nlambda * a.nosm * a.npsm * a.nslice;
    if (nF > slots)
      return set_err(PTY_EINVAL, "instance %d: %d exit-wave fields exceed %d for Np=%d", idx, nF, slots, a.Np);
    if (nF > 32) return set_err(PTY_EINVAL, "instance %d: %d exit-wave fields exceed 32", idx, nF);
    if ((a.engine == PTY_MPIE || a.engine == PTY_MULTIPIE) && a.momentum_mask &&
        (!a.objectBuffer || !a.objectMomentum || !a.probeBuffer || !a.probeMomentum))
      return set_err(PTY_EINVAL, "instance %d: momentum needs buffer and momentum arrays", idx);
  }
  return PTY_OK;
}

}  // namespace

template <int NP>
static int launch_propagate(float2* data, int nf, int prop, int inverse, const float2* tf, hipStream_t st) {
  const int lines = nf * NP;
  const int per_block = NWAVE * Line<NP>::SUB;
  const int blocks = (lines + per_block - 1) / per_block;
  if (prop == PTY_ASP) {
    pty_rows_fwd_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf);
    pty_asp_cols_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf, tf, inverse);
    pty_rows_inv_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf);
  } else {
    const int cbm = prop == PTY_FRAUNHOFER ? 1 : 0;
    const int inv = inverse ? LN_INVERSE : 0;
    pty_lines_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf, inv | (cbm ? LN_PRE_CB : 0));
    pty_lines_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf, inv | LN_COLS | LN_SCALE | (cbm ? LN_POST_CB : 0));
  }
  HIP_TRY(hipGetLastError());
  return PTY_OK;
}

template <int NP>
static int launch_fft2(float2* data, int nf, int sign, hipStream_t st) {
  const int lines = nf * NP;
  const int per_block = NWAVE * Line<NP>::SUB;
  const int blocks = (lines + per_block - 1) / per_block;
  const int inv = sign > 0 ? LN_INVERSE : 0;
  pty_lines_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf, inv);
  pty_lines_kernel<NP><<<blocks, NTHREADS, 0, st>>>(data, nf, inv | LN_COLS);
  HIP_TRY(hipGetLastError());
  return PTY_OK;
}

extern "C" {

int pty_internal_set_error(int code, const char* msg) __attribute__((visibility("hidden")));
int pty_internal_set_error(int code, const char* msg) { return set_err(code, "%s", msg); }

int pty_abi_version(void) { return PTY_ABI_VERSION; }

const char* pty_last_error(void) { return g_err.c_str(); }

// HIP_TRY inside pty_plan_create: record the error and leave through the
// function's `fail` label, which frees the half-built plan
#define PLAN_TRY(expr)                                                                       \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      rc = set_err(PTY_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_));              \
      goto fail;                                                                             \
    }                                                                                        \
  } while (0)

int pty_plan_create(pty_plan** out, const pty_instance* insts, int32_t ninst, void* hip_stream) {
  if (!out || !insts || ninst < 1) return set_err(PTY_EINVAL, "pty_plan_create: bad arguments");
  *out = nullptr;
  int rc = ensure_twiddles();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)hip_stream;
  pty_plan* p = new pty_plan();
  {
    const hipError_t e = hipGetDevice(&p->device);
    if (e != hipSuccess) {
      delete p;
      return set_err(PTY_EDEVICE, "hipGetDevice failed: %s", hipGetErrorString(e));
    }
  }
  p->ninst = ninst;
  p->h_insts.resize(ninst);
  std::vector<std::vector<int2>> cmaps;  // per launch class
  for (int i = 0; i < ninst; ++i) {
    const pty_instance& a = insts[i];
    const bool batched = joins_batch(a, ninst);
    const int nt = batched ? BATCH_NT : nthreads_rt(a.Np);
    rc = validate(a, i, nt);
    if (rc) {
      free_plan(p);
      delete p;
      return rc;
    }
    DevInst& d = p->h_insts[i];
    memset(&d, 0, sizeof d);
    d.a = a;
    int G = a.workgroups > 0 ? a.workgroups : auto_workgroups(a, ninst, nt);
    if (batched) {
      // a batched cluster runs on one XCD
      int xcap = 0;
      if ((rc = xcd_capacity(family_of(a), &xcap))) goto fail;
      if (a.workgroups > xcap) {
        rc = set_err(PTY_EINVAL, "instance %d: %d workgroups exceed the %d of one XCD that a batched cluster may use", i,
                     a.workgroups, xcap);
        goto fail;
      }
      if (G > xcap) G = xcap;
    }
    if (G > a.Np) G = a.Np;
    if (G < 1) G = 1;
    while (a.Np % G) --G;
    d.G = G;
    const int np = a.Np;
    const size_t np2 = (size_t)np * np;
    if (a.engine == PTY_E3PIE) {
      d.nF = a.nlambda * a.npsm;
      d.fs = 1;
    } else {
      d.nF = a.nlambda * a.nosm * a.npsm * a.nslice;
      d.fs = a.nslice;
    }
    d.nFp = 1;
    while (d.nFp < d.nF) d.nFp *= 2;
    if (a.engine != PTY_E3PIE) {
      // the field sums of the object and probe updates (ePIE.py:128-160,
      // mPIE.py:226-270, multiPIE.py:153-205, qNewton.py:116-140)
      const bool full = a.engine == PTY_EPIE || a.engine == PTY_QNEWTON || a.engine == PTY_MQNEWTON;
      for (int f = 0; f < d.nF; ++f) {
        const FieldIdx x = field_idx(f, a.nosm, a.npsm, a.nslice);
        for (int g = 0; g < d.nF; ++g) {
          const FieldIdx y = field_idx(g, a.nosm, a.npsm, a.nslice);
          if (full ? y.io == x.io : y.fo == x.fo) d.osel[f] |= 1u << g;
          if (full ? y.ip == x.ip : (y.ip == x.ip && y.is == x.is && (a.engine == PTY_MULTIPIE || y.il == x.il)))
            d.psel[f] |= 1u << g;
        }
      }
    }
    d.nO = a.nlambda * a.nosm * a.nslice;
    d.nPr = a.nlambda * a.npsm * a.nslice;
    d.fl = 0;
    if ((a.engine == PTY_EPIE || a.engine == PTY_MPIE || a.engine == PTY_MULTIPIE) && d.nO == 1 && d.nF == d.nPr) {
      bool ok = d.osel[0] == (d.nF >= 32 ? ~0u : (1u << d.nF) - 1u);
      for (int f = 0; f < d.nF; ++f) ok = ok && d.psel[f] == (1u << f);
      d.fl = ok ? 1 : 0;
    }
    d.omAlphaO = (float)(1.0 - (double)a.alphaObject);
    d.omAlphaP = (float)(1.0 - (double)a.alphaProbe);
    // positions must keep every patch inside the object
    std::vector<int32_t> pos((size_t)a.F * 2);
    PLAN_TRY(hipMemcpy(pos.data(), a.positions, pos.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int f = 0; f < a.F; ++f) {
      const int r = pos[2 * f], c = pos[2 * f + 1];
      if (r < 0 || c < 0 || r + np > a.No || c + np > a.No) {
        free_plan(p);
        delete p;
        return set_err(PTY_EINVAL, "instance %d: position %d (%d, %d) puts the %d-pixel patch outside the %d-pixel object",
                       i, f, r, c, np, a.No);
      }
    }
    std::vector<int32_t> ord((size_t)a.F * a.n_iter);
    PLAN_TRY(hipMemcpy(ord.data(), a.order, ord.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < ord.size(); ++q) {
      if (ord[q] < 0 || ord[q] >= a.F) {
        free_plan(p);
        delete p;
        return set_err(PTY_EINVAL, "instance %d: order entry %zu = %d out of range", i, q, ord[q]);
      }
    }
    void* mem = nullptr;
    if ((rc = dalloc(p, &mem, (size_t)d.nF * np2 * sizeof(float2)))) goto fail;
    d.W = (float2*)mem;
    if (a.engine == PTY_E3PIE) {
      if ((rc = dalloc(p, &mem, (size_t)a.nslice * d.nF * np2 * sizeof(float2)))) goto fail;
      d.ESW = (float2*)mem;
    }
    if (a.engine == PTY_QNEWTON || a.engine == PTY_MQNEWTON || a.pc_direction) {
      if ((rc = dalloc(p, &mem, (size_t)d.nO * np2 * sizeof(float2)))) goto fail;
      d.OP = (float2*)mem;
    }
    if (a.pc_direction) {
      if ((rc = dalloc(p, &mem, (size_t)G * PC_TERMS * sizeof(double)))) goto fail;
      d.pcpart = (double*)mem;
      if (a.pc_radius >= 2) {
        if ((rc = dalloc(p, &mem, (size_t)2 * np2 * sizeof(float2)))) goto fail;
        d.PCW = (float2*)mem;
        const size_t nw = (size_t)(2 * a.pc_radius + 1);
        if ((rc = dalloc(p, &mem, nw * nw * sizeof(double)))) goto fail;
        d.pcc = (double*)mem;
      }
    }
    if (a.engine == PTY_MQNEWTON) {
      if ((rc = dalloc(p, &mem, (size_t)2 * G * sizeof(double)))) goto fail;
      d.adampart = (double*)mem;
      // the second moments start at zero (a fresh engine's momentum copy,
      // mqNewton.py:68-79) and persist across the plan's launches
      if ((rc = dalloc(p, &mem, 4 * sizeof(double)))) goto fail;
      PLAN_TRY(hipMemsetAsync(mem, 0, 4 * sizeof(double), st));
      d.adamv = (double*)mem;
    }
    if ((rc = dalloc(p, &mem, CTRL_WORDS * sizeof(unsigned)))) goto fail;
    d.ctrl = (unsigned*)mem;
    if ((rc = dalloc(p, &mem, (size_t)4 * SLOT_WORDS * G * sizeof(unsigned)))) goto fail;
    d.mx = (unsigned*)mem;
    if ((rc = dalloc(p, &mem, (size_t)G * sizeof(unsigned)))) goto fail;
    d.bar = (unsigned*)mem;
    if ((rc = dalloc(p, &mem, (size_t)2 * a.F * G * sizeof(double)))) goto fail;
    d.errpart = (double*)mem;
    if (prop_asp(a.propagator)) {
      d.Ik = a.ptychogram;  // natural order: the constraint runs in the row pass
    } else {
      if ((rc = dalloc(p, &mem, (size_t)a.F * np2 * sizeof(float)))) goto fail;
      pty_permute_frames_kernel<<<1024, 256, 0, st>>>(a.ptychogram, (float*)mem, a.F, np, log2_np(np));
      PLAN_TRY(hipGetLastError());
      d.Ik = (const float*)mem;
    }
    d.fpl = d.nF;
    if (a.propagator == PTY_FRESNEL) {
      d.Q = (const float2*)a.tf;  // natural order, read by the row passes
    } else if (a.tf) {
      const bool per_lambda = a.propagator == PTY_POLYCHROME_ASP || a.propagator == PTY_SCALED_POLYCHROME_ASP;
      const int ntf = per_lambda ? a.nlambda : 1;
      if (per_lambda) d.fpl = d.nF / a.nlambda;
      if ((rc = dalloc(p, &mem, (size_t)ntf * np2 * sizeof(float2)))) goto fail;
      for (int t = 0; t < ntf; ++t)
        pty_permute_tf_kernel<<<256, 256, 0, st>>>((const float2*)a.tf + (size_t)t * np2, (float2*)mem + (size_t)t * np2,
                                                   np, log2_np(np));
      PLAN_TRY(hipGetLastError());
      d.HT = (const float2*)mem;
      if (a.propagator == PTY_SCALED_ASP || a.propagator == PTY_SCALED_POLYCHROME_ASP)
        d.Q = (const float2*)a.tf + (size_t)ntf * np2;  // c Q1 per wavelength, natural order
      if (prop_asp(a.propagator) && a.engine != PTY_E3PIE) {
        if ((rc = dalloc(p, &mem, (size_t)ntf * np))) goto fail;
        pty_live_columns_kernel<<<(ntf * np + 255) / 256, 256, 0, st>>>(d.HT, (unsigned char*)mem, np, ntf);
        PLAN_TRY(hipGetLastError());
        d.live = (const unsigned char*)mem;
        // column width of the kernel class the instance runs in
        const int cw = batched ? col_width<BATCH_CW>(np) : col_width<0>(np);
        if ((rc = dalloc(p, &mem, (size_t)ntf * (np / cw) * sizeof(short)))) goto fail;
        d.lgrp = (const short*)mem;
        if ((rc = dalloc(p, &mem, (size_t)(ntf + 1) * sizeof(int)))) goto fail;
        d.nlive = (const int*)mem;
        // packed exchange rows for the PIE families in batched launches;
        // qNewton's update pass writes full real-space rows into W that its
        // probe pass reads.  (The per-size kernels keep full rows: there the
        // column table's load before every row access lengthened the
        // position, C1 29.9 -> 30.6 us, while the batched launch gained 22 %.)
        const bool packed = a.engine != PTY_QNEWTON && a.engine != PTY_MQNEWTON;
        if (packed) {
          if ((rc = dalloc(p, &mem, (size_t)ntf * np * sizeof(short)))) goto fail;
          d.cix = (const short*)mem;
        }
        pty_live_groups_kernel<<<1, 64, 0, st>>>(d.live, (short*)d.lgrp, (int*)d.nlive, (short*)d.cix, np, cw, ntf);
        PLAN_TRY(hipGetLastError());
        if (packed) {
          int most = 0;
          PLAN_TRY(hipStreamSynchronize(st));
          PLAN_TRY(hipMemcpy(&most, d.nlive, sizeof(int), hipMemcpyDeviceToHost));
          d.wst = most * cw;
        }
      }
    }
    if (a.n_iter > p->max_iter) p->max_iter = a.n_iter;
    int ci = -1;
    const int fam = family_of(a);
    const int key = batched ? 0 : a.Np;
    for (size_t c = 0; c < p->classes.size(); ++c)
      if (p->classes[c].np == key && p->classes[c].fam == fam) ci = (int)c;
    if (ci < 0) {
      p->classes.emplace_back();
      p->classes.back().np = key;
      p->classes.back().fam = fam;
      p->classes.back().nthreads = nt;
      p->use_graph = p->use_graph || batched;
      cmaps.emplace_back();
      ci = (int)p->classes.size() - 1;
    }
    if (G > 1) p->classes[ci].cooperative = true;
    for (int g = 0; g < G; ++g) {
      d.gsize[cmaps[ci].size() & 7] += 1;
      cmaps[ci].push_back(make_int2(i, g));
    }
    for (int x = 0; x < 8; ++x) d.ngroups += d.gsize[x] > 0;
  }
  // Batched classes: each cluster goes whole onto one XCD, the largest
  // clusters first, each to the least-loaded XCD; cluster member g of an
  // instance on XCD x in slot s takes block 8 s + x.  Unused blocks get
  // instance -1 and exit at once.
  for (size_t c = 0; c < p->classes.size(); ++c) {
    if (p->classes[c].np != 0) continue;
    std::vector<int> ids;
    for (const int2& e : cmaps[c])
      if (e.y == 0) ids.push_back(e.x);
    std::stable_sort(ids.begin(), ids.end(), [&](int u, int v) { return p->h_insts[u].G > p->h_insts[v].G; });
    int xcap = 0;
    if ((rc = xcd_capacity(p->classes[c].fam, &xcap))) goto fail;
    int used[N_XCD] = {0};
    std::vector<int2> m;
    for (const int i : ids) {
      DevInst& d = p->h_insts[i];
      int x = 0;
      for (int y = 1; y < N_XCD; ++y)
        if (used[y] < used[x]) x = y;
      if (used[x] + d.G > xcap) {
        rc = set_err(PTY_EINVAL, "batched launch of family %d: the clusters do not fit on %d XCDs of %d workgroups",
                     p->classes[c].fam, N_XCD, xcap);
        goto fail;
      }
      for (int g = 0; g < d.G; ++g) {
        const size_t b = (size_t)N_XCD * (used[x] + g) + x;
        if (m.size() <= b) m.resize(b + 1, make_int2(-1, 0));
        m[b] = make_int2(i, g);
      }
      used[x] += d.G;
    }
    int nslots = 0;
    for (int y = 0; y < N_XCD; ++y) nslots = used[y] > nslots ? used[y] : nslots;
    m.resize((size_t)N_XCD * nslots, make_int2(-1, 0));  // every XCD receives nslots blocks
    p->classes[c].nslots = nslots;
    cmaps[c].swap(m);
  }
  {
    void* mem = nullptr;
    if ((rc = dalloc(p, &mem, sizeof(DevInst) * ninst))) goto fail;
    p->d_insts = (DevInst*)mem;
    p->nclaim = (int)p->classes.size() * N_XCD * 32;
    if ((rc = dalloc(p, &mem, sizeof(unsigned) * p->nclaim))) goto fail;
    p->d_claims = (unsigned*)mem;
    PLAN_TRY(hipMemsetAsync(p->d_claims, 0, sizeof(unsigned) * p->nclaim, st));
    for (size_t c = 0; c < p->classes.size(); ++c) p->classes[c].d_claim = p->d_claims + c * N_XCD * 32;
    PLAN_TRY(hipMemcpyAsync(p->d_insts, p->h_insts.data(), sizeof(DevInst) * ninst, hipMemcpyHostToDevice, st));
    int cus = 0, coop = 0;
    PLAN_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p->device));
    PLAN_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, p->device));
    p->nblocks = 0;
    for (size_t c = 0; c < p->classes.size(); ++c) {
      LaunchClass& lc = p->classes[c];
      lc.nblocks = (int)cmaps[c].size();
      for (const int2& e : cmaps[c]) p->nblocks += e.x >= 0;
      if ((rc = dalloc(p, &mem, sizeof(int2) * cmaps[c].size()))) goto fail;
      lc.d_bmap = (int2*)mem;
      PLAN_TRY(hipMemcpyAsync(lc.d_bmap, cmaps[c].data(), sizeof(int2) * cmaps[c].size(), hipMemcpyHostToDevice, st));
      if (lc.cooperative) {
        // clusters spin on each other: the whole launch must be co-resident
        int per_cu = 0;
        PLAN_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, iterate_kernel_ptr(lc.np, lc.fam), lc.nthreads, 0));
        if (!coop) {
          rc = set_err(PTY_EDEVICE, "device does not support cooperative launches");
          goto fail;
        }
        if (lc.nblocks > per_cu * cus) {
          rc = set_err(PTY_EINVAL, "%d workgroups for %s%d exceed the %d that can be co-resident", lc.nblocks,
                       lc.np ? "Np=" : "the batched launch of family ", lc.np ? lc.np : lc.fam, per_cu * cus);
          goto fail;
        }
      }
    }
    PLAN_TRY(hipStreamSynchronize(st));
  }
  *out = p;
  return PTY_OK;
fail:
  (void)hipStreamSynchronize(st);
  free_plan(p);
  delete p;
  return rc;
}

static int plan_launch(pty_plan* p, int32_t it_begin, int32_t it_end, int32_t jb, int32_t je, void* hip_stream);

int pty_plan_run(pty_plan* p, int32_t it_begin, int32_t it_end, void* hip_stream) {
  return plan_launch(p, it_begin, it_end, 0, 0x7fffffff, hip_stream);
}

int pty_plan_run_positions(pty_plan* p, int32_t it, int32_t j_begin, int32_t j_end, void* hip_stream) {
  if (j_begin < 0 || j_end < j_begin) return set_err(PTY_EINVAL, "pty_plan_run_positions: bad position range");
  if (j_end == j_begin) return PTY_OK;
  return plan_launch(p, it, it + 1, j_begin, j_end, hip_stream);
}

// Per-iteration graph of a plan: the control-block reset (a fill kernel, no
// memset node), then the class kernels in sequence.  Kernel nodes are plain
// launches: the co-residency that a cooperative launch would verify is
// checked once at plan creation.  Each replay rewrites the iteration
// arguments of the kernel nodes, then launches the instantiated graph.
static int graph_build(pty_plan* p) {
  HIP_TRY(hipGraphCreate(&p->graph, 0));
  const DevInst* di = p->d_insts;
  int zero = 0;
  hipKernelNodeParams kp = {};
  unsigned* claims = p->d_claims;
  int nclaim = p->nclaim;
  void* rargs[] = {(void*)&di, (void*)&claims, (void*)&nclaim};
  kp.func = (void*)pty_reset_ctrl_kernel;
  kp.gridDim = dim3(p->ninst);
  kp.blockDim = dim3(64);
  kp.sharedMemBytes = 0;
  kp.kernelParams = rargs;
  kp.extra = nullptr;
  hipGraphNode_t prev = nullptr;
  HIP_TRY(hipGraphAddKernelNode(&prev, p->graph, nullptr, 0, &kp));
  for (LaunchClass& lc : p->classes) {
    const int2* bm = lc.d_bmap;
    unsigned* cl = lc.d_claim;
    void* args[] = {(void*)&di, (void*)&bm, (void*)&zero, (void*)&zero, (void*)&zero, (void*)&zero, (void*)&cl,
                    (void*)&lc.nslots};
    kp.func = (void*)iterate_kernel_ptr(lc.np, lc.fam);
    kp.gridDim = dim3(lc.nblocks);
    kp.blockDim = dim3(lc.nthreads);
    kp.kernelParams = args;
    HIP_TRY(hipGraphAddKernelNode(&lc.node, p->graph, &prev, 1, &kp));
    prev = lc.node;
  }
  HIP_TRY(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
  return PTY_OK;
}

static int graph_replay(pty_plan* p, int it0, int it1, int jb, int je, hipStream_t st) {
  if (!p->exec) {
    const int rc = graph_build(p);
    if (rc) return rc;
  }
  const DevInst* di = p->d_insts;
  for (LaunchClass& lc : p->classes) {
    const int2* bm = lc.d_bmap;
    unsigned* cl = lc.d_claim;
    void* args[] = {(void*)&di, (void*)&bm, (void*)&it0, (void*)&it1, (void*)&jb, (void*)&je, (void*)&cl, (void*)&lc.nslots};
    hipKernelNodeParams kp = {};
    kp.func = (void*)iterate_kernel_ptr(lc.np, lc.fam);
    kp.gridDim = dim3(lc.nblocks);
    kp.blockDim = dim3(lc.nthreads);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    HIP_TRY(hipGraphExecKernelNodeSetParams(p->exec, lc.node, &kp));
  }
  HIP_TRY(hipGraphLaunch(p->exec, st));
  return PTY_OK;
}

static int plan_launch(pty_plan* p, int32_t it_begin, int32_t it_end, int32_t jb, int32_t je, void* hip_stream) {
  if (!p) return set_err(PTY_EINVAL, "pty_plan_run: null plan");
  if (it_begin < 0 || it_end < it_begin) return set_err(PTY_EINVAL, "pty_plan_run: bad iteration range");
  if (it_end > p->max_iter) return set_err(PTY_EINVAL, "pty_plan_run: it_end %d > n_iter %d", it_end, p->max_iter);
  if (it_end == it_begin) return PTY_OK;
  hipStream_t st = (hipStream_t)hip_stream;
  if (p->use_graph) {
    // one replay per iteration
    for (int it = it_begin; it < it_end; ++it) {
      const int rc = graph_replay(p, it, it + 1, it == it_begin ? jb : 0, it == it_end - 1 ? je : 0x7fffffff, st);
      if (rc) return rc;
    }
    return PTY_OK;
  }
  pty_reset_ctrl_kernel<<<p->ninst, 64, 0, st>>>(p->d_insts, p->d_claims, p->nclaim);
  HIP_TRY(hipGetLastError());
  const DevInst* di = p->d_insts;
  int a0 = it_begin, a1 = it_end, b0 = jb, b1 = je;
  for (LaunchClass& lc : p->classes) {
    const int2* bm = lc.d_bmap;
    unsigned* cl = lc.d_claim;
    void* args[] = {(void*)&di, (void*)&bm, (void*)&a0, (void*)&a1, (void*)&b0, (void*)&b1, (void*)&cl, (void*)&lc.nslots};
    if (lc.cooperative)
      HIP_TRY(hipLaunchCooperativeKernel(iterate_kernel_ptr(lc.np, lc.fam), dim3(lc.nblocks), dim3(lc.nthreads), args, 0, st));
    else
      HIP_TRY(hipLaunchKernel(iterate_kernel_ptr(lc.np, lc.fam), dim3(lc.nblocks), dim3(lc.nthreads), args, 0, st));
  }
  return PTY_OK;
}

int pty_plan_use_graph(pty_plan* p, int32_t enable) {
  if (!p) return set_err(PTY_EINVAL, "pty_plan_use_graph: null plan");
  p->use_graph = enable != 0;
  return PTY_OK;
}

int pty_coresident_workgroups(int32_t Np, int32_t engine, int32_t propagator, int32_t batched, int32_t* count) {
  if (!count || log2_np(Np) < 0 || engine < PTY_EPIE || engine > PTY_MQNEWTON || propagator < PTY_FRAUNHOFER ||
      propagator > PTY_SCALED_POLYCHROME_ASP)
    return set_err(PTY_EINVAL, "pty_coresident_workgroups: bad arguments");
  pty_instance a;
  memset(&a, 0, sizeof a);
  a.Np = Np;
  a.engine = engine;
  a.propagator = propagator;
  a.flags = batched ? PTY_FLAG_BATCHED : 0;
  const bool b = joins_batch(a, 1);
  const int nt = b ? BATCH_NT : nthreads_rt(Np);
  int dev = 0, cus = 0, per_cu = 0;
  HIP_TRY(hipGetDevice(&dev));
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, iterate_kernel_ptr(b ? 0 : Np, family_of(a)), nt, 0));
  *count = per_cu * cus;
  return PTY_OK;
}

int pty_plan_check(pty_plan* p, void* hip_stream) {
  if (!p) return set_err(PTY_EINVAL, "pty_plan_check: null plan");
  hipStream_t st = (hipStream_t)hip_stream;
  HIP_TRY(hipStreamSynchronize(st));
  for (size_t i = 0; i < p->h_insts.size(); ++i) {
    unsigned w[4] = {0u, 0u, 0u, 0u};
    HIP_TRY(hipMemcpy(w, p->h_insts[i].ctrl, sizeof w, hipMemcpyDeviceToHost));
    if (w[CTRL_BADPOS]) return set_err(PTY_EINVAL, "instance %zu: a position puts the patch outside the object", i);
    if (w[CTRL_PLACE])
      return set_err(PTY_EDEVICE, "instance %zu: a workgroup of its XCD-local cluster ran on another XCD", i);
    if (w[CTRL_FAIL]) return set_err(PTY_ETIMEDOUT, "instance %zu: a cluster barrier timed out", i);
  }
  return PTY_OK;
}

int pty_plan_workgroups(const pty_plan* p, int32_t* total) {
  if (!p || !total) return set_err(PTY_EINVAL, "pty_plan_workgroups: bad arguments");
  *total = p->nblocks;
  return PTY_OK;
}

int pty_plan_destroy(pty_plan* p) {
  if (!p) return PTY_OK;
  (void)hipDeviceSynchronize();
  free_plan(p);
  delete p;
  return PTY_OK;
}

int pty_propagate(void* fields, int32_t nfields, int32_t Np, int32_t propagator, int32_t inverse, const void* tf,
                  void* hip_stream) {
  if (!fields || nfields < 1) return set_err(PTY_EINVAL, "pty_propagate: bad arguments");
  if (propagator < PTY_FRAUNHOFER || propagator > PTY_ASP) return set_err(PTY_EINVAL, "pty_propagate: bad propagator");
  if (propagator == PTY_ASP && !tf) return set_err(PTY_EINVAL, "pty_propagate: ASP needs a transfer function");
  int rc = ensure_twiddles();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)hip_stream;
  float2* d = (float2*)fields;
  const float2* h = (const float2*)tf;
  switch (Np) {
    case 16: return launch_propagate<16>(d, nfields, propagator, inverse, h, st);
    case 32: return launch_propagate<32>(d, nfields, propagator, inverse, h, st);
    case 64: return launch_propagate<64>(d, nfields, propagator, inverse, h, st);
    case 128: return launch_propagate<128>(d, nfields, propagator, inverse, h, st);
    case 256: return launch_propagate<256>(d, nfields, propagator, inverse, h, st);
    case 512: return launch_propagate<512>(d, nfields, propagator, inverse, h, st);
  }
  return set_err(PTY_EINVAL, "pty_propagate: Np=%d unsupported (powers of two 16..512)", Np);
}

int pty_fft2(void* fields, int32_t nfields, int32_t Np, int32_t sign, void* hip_stream) {
  if (!fields || nfields < 1 || (sign != 1 && sign != -1)) return set_err(PTY_EINVAL, "pty_fft2: bad arguments");
  int rc = ensure_twiddles();
  if (rc) return rc;
  hipStream_t st = (hipStream_t)hip_stream;
  float2* d = (float2*)fields;
  switch (Np) {
    case 16: return launch_fft2<16>(d, nfields, sign, st);
    case 32: return launch_fft2<32>(d, nfields, sign, st);
    case 64: return launch_fft2<64>(d, nfields, sign, st);
    case 128: return launch_fft2<128>(d, nfields, sign, st);
    case 256: return launch_fft2<256>(d, nfields, sign, st);
    case 512: return launch_fft2<512>(d, nfields, sign, st);
  }
  return set_err(PTY_EINVAL, "pty_fft2: Np=%d unsupported (powers of two 16..512)", Np);
}

}  // extern "C"
