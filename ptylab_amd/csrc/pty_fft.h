// Wavefront-resident FFT building blocks for gfx950 (wave64).
//
// A length-N line lives in registers of a group of LPF lanes: element
// n = l * LPF + t sits in register l of the lane whose in-line index is t
// (L = N / LPF registers per lane).  Consecutive in-line indices are LS
// lanes apart in the wave (LS = 1: a line occupies adjacent lanes; LS = CW:
// CW lines interleaved lane by lane, as the column passes use).
// Butterflies whose span is a multiple of LPF are done inside a lane; smaller
// spans exchange values between lanes (mask span * LS) with DPP lane moves
// and the permlane swaps (xlane) -- the cross-lane form, used when a lane
// holds fewer than 4 values of the line.  With 4 or more values per lane the
// transposed form below runs every butterfly inside a lane and moves the line
// between register layouts through a wave-private LDS scratch.  No workgroup
// barrier is needed inside a transform.
//
// fft_fwd: radix-2 decimation in frequency, natural order in, bit-reversed
//          order out, kernel exp(-2 pi i nk / N), unnormalised.
// fft_inv: radix-2 decimation in time, bit-reversed order in, natural order
//          out, kernel exp(+2 pi i nk / N), unnormalised.
// Running fft_inv(fft_fwd(x)) gives N * x, so a forward transform whose
// output is consumed in bit-reversed order never needs a permutation.
//
// `tw` holds exp(-2 pi i j / N), j < N / 2, computed in double on the host
// and rounded to float.
#pragma once
#include <hip/hip_runtime.h>

namespace pty {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }

// Complex product with fused multiply-adds (two roundings per component
// instead of three); used inside the transforms only.
__device__ __forceinline__ float2 cmul_f(float2 a, float2 b) {
  return make_float2(__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x));
}

// The value of lane (lane ^ m) for m = 1, 2, ..., 32, without LDS: DPP lane
// moves inside a row of 16 lanes, the gfx950 permlane swaps across rows.
// m must be a constant once the caller's loops are unrolled (every branch
// below but one folds away).  Exact data movement: results are the same as
// with ds_bpermute.
//   m = 1, 2: quad_perm [1,0,3,2] / [2,3,0,1];
//   m = 4: half-row mirror (lane i of 8 reads 7 - i = i ^ 7), then
//          quad_perm [3,2,1,0] (i ^ 3): i ^ 7 ^ 3 = i ^ 4;
//   m = 8: row mirror (15 - i = i ^ 15), then half-row mirror (i ^ 7);
//   m = 16: v_permlane16_swap of x with itself leaves row r's partner row
//           (r ^ 1) in the second result for even rows, in the first for odd;
//   m = 32: v_permlane32_swap likewise for the two halves of the wave.
__device__ __forceinline__ int xlane_i(int x, int m) {
  int r;
  if (m == 1) {
    r = __builtin_amdgcn_mov_dpp(x, 0xB1, 0xF, 0xF, false);
  } else if (m == 2) {
    r = __builtin_amdgcn_mov_dpp(x, 0x4E, 0xF, 0xF, false);
  } else if (m == 4) {
    r = __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
  } else if (m == 8) {
    r = __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(x, 0x140, 0xF, 0xF, false), 0x141, 0xF, 0xF, false);
  } else if (m == 16) {
    const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    r = (__lane_id() & 16) ? (int)p[0] : (int)p[1];
  } else {
    const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    r = (__lane_id() & 32) ? (int)p[0] : (int)p[1];
  }
  return r;
}
__device__ __forceinline__ float xlane(float v, int m) { return __int_as_float(xlane_i(__float_as_int(v), m)); }
__device__ __forceinline__ float2 xlane2(float2 v, int m) { return make_float2(xlane(v.x, m), xlane(v.y, m)); }
__device__ __forceinline__ double xlane_d(double v, int m) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)xlane_i((int)(unsigned)u, m), hi = (unsigned)xlane_i((int)(unsigned)(u >> 32), m);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <int N>
struct Line {
  static constexpr int LPF = N >= 64 ? 64 : N;  // lanes per line (row layout)
  static constexpr int L = N / LPF;             // values per lane
  static constexpr int SUB = 64 / LPF;          // lines per wave
  static constexpr int LOG2 = N == 16 ? 4 : N == 32 ? 5 : N == 64 ? 6 : N == 128 ? 7 : N == 256 ? 8 : 9;
};

template <int N, int LPF, int LS>
__device__ __forceinline__ void fft_fwd_x(float2 (&v)[N / LPF], const float2* __restrict__ tw_, int t) {
  constexpr int L = N / LPF;
  // Opaque zero offset: keeps the compiler from hoisting the (loop-invariant)
  // twiddle loads out of the caller's loops into registers.
  unsigned off = 0;
  __asm__ volatile("" : "+s"(off));
  const float2* __restrict__ tw = tw_ + off;
#pragma unroll
  for (int sr = L / 2; sr >= 1; sr >>= 1) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if ((l & sr) == 0) {
        const int i = (l & (sr - 1)) * LPF + t;
        const float2 w = tw[i * (L / (2 * sr))];
        const float2 a = v[l], b = v[l + sr];
        v[l] = cadd(a, b);
        v[l + sr] = cmul_f(csub(a, b), w);
      }
    }
  }
  // Cross-lane stages without divergence: the lower lane of a pair keeps
  // x + p, the upper one (p - x) w; both are (p + sgn x) w' with sgn = -1,
  // w' = w in the upper lane and sgn = 1, w' = 1 in the lower one.
#pragma unroll
  for (int s = LPF / 2; s >= 1; s >>= 1) {
    const bool up = (t & s) != 0;
    const float2 w0 = tw[(t & (s - 1)) * (N / (2 * s))];
    const float2 w = up ? w0 : make_float2(1.f, 0.f);
    const float sgn = up ? -1.f : 1.f;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const float2 x = v[l];
      const float2 p = xlane2(x, s * LS);
      const float2 y = make_float2(__builtin_fmaf(sgn, x.x, p.x), __builtin_fmaf(sgn, x.y, p.y));
      v[l] = cmul_f(y, w);
    }
  }
}

template <int N, int LPF, int LS>
__device__ __forceinline__ void fft_inv_x(float2 (&v)[N / LPF], const float2* __restrict__ tw_, int t) {
  constexpr int L = N / LPF;
  // Opaque zero offset: keeps the compiler from hoisting the (loop-invariant)
  // twiddle loads out of the caller's loops into registers.
  unsigned off = 0;
  __asm__ volatile("" : "+s"(off));
  const float2* __restrict__ tw = tw_ + off;
  // cross-lane stages without divergence: the upper lane of a pair first
  // multiplies by w (the lower one by 1), then keeps p - x, the lower x + p;
  // both are p + sgn x
#pragma unroll
  for (int s = 1; s < LPF; s <<= 1) {
    const bool up = (t & s) != 0;
    const float2 w0 = conjf2(tw[(t & (s - 1)) * (N / (2 * s))]);
    const float2 w = up ? w0 : make_float2(1.f, 0.f);
    const float sgn = up ? -1.f : 1.f;
#pragma unroll
    for (int l = 0; l < L; ++l) {
      const float2 x = cmul_f(v[l], w);
      const float2 p = xlane2(x, s * LS);
      v[l] = make_float2(__builtin_fmaf(sgn, x.x, p.x), __builtin_fmaf(sgn, x.y, p.y));
    }
  }
#pragma unroll
  for (int sr = 1; sr < L; sr <<= 1) {
#pragma unroll
    for (int l = 0; l < L; ++l) {
      if ((l & sr) == 0) {
        const int i = (l & (sr - 1)) * LPF + t;
        const float2 w = conjf2(tw[i * (L / (2 * sr))]);
        const float2 a = v[l];
        const float2 b = cmul_f(v[l + sr], w);
        v[l] = cadd(a, b);
        v[l + sr] = csub(a, b);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Transposed form: every butterfly inside a lane.
// ---------------------------------------------------------------------------
// With L >= 4 values per lane the same radix-2 stages run without any
// cross-lane arithmetic: the line is rewritten through a wave-private LDS
// scratch so that the lg = log2(L) bits of n a group of stages pairs on are
// register bits, and every butterfly is the in-lane one.  Layout k keeps bits
// [k, k + lg) of n in the register index and the other bits in the lane:
//   n(l, t; k) = ((t >> k) << (k + lg)) | (l << k) | (t & (2^k - 1)).
// Layout k0 = log2(N) - lg is the caller's (n = l * P + t).  The forward
// transform (decimation in frequency, bit 7 ... bit 0 for N = 256) visits
// k0, k0 - lg, ..., 0 and returns to k0; the inverse visits them in reverse.
// Each stage computes exactly what the cross-lane form computes (same
// twiddle, same operations), so the results are the same bit for bit (up to
// the sign of exact zeros: the cross-lane form multiplies the lower output
// by 1 + 0i).
//
// Scratch addresses: element n of line c (the LS lines of a wave, lanes
// t * LS + c) lives at c * STRIDE + n + cc * ((n >> s) & msk), with (s, msk,
// cc) chosen per transpose so that the map is one-to-one and both the write
// layout and the read layout are free of bank conflicts (a 64-bit access is
// served 16 lanes at a time from 16 distinct 8-byte banks); the tables below
// were found by exhaustive search over that family and are re-checked by
// static_assert.  The address of n = (t part) | (l part) splits into a per-lane
// base plus a constant offset per register, so every LDS access uses an
// immediate offset.

constexpr int ilog2c(int x) { return x <= 1 ? 0 : 1 + ilog2c(x >> 1); }

struct TPad {
  int s, msk, cc, size, stride;  // padding n + cc * ((n >> s) & msk); elements of one line; line stride
};

// Padding of the transpose between layouts k1 and k2 (symmetric), for a
// line of N points on P lanes, LS lines per wave.
struct TPadEntry {
  int N, P, LS, a, b;  // configuration and layout pair (a < b)
  TPad pad;
};
constexpr TPadEntry kTPad[] = {
    {64, 16, 4, 0, 2, {2, 7, 5, 99, 100}},      {64, 16, 4, 0, 4, {2, 7, 5, 99, 100}},
    {64, 16, 4, 2, 4, {0, 0, 0, 64, 68}},       {64, 8, 8, 0, 3, {3, 7, 1, 71, 74}},
    {128, 32, 2, 0, 1, {2, 31, 1, 159, 168}},   {128, 32, 2, 0, 5, {3, 15, 1, 143, 152}},
    {128, 32, 2, 1, 3, {3, 15, 2, 158, 168}},   {128, 32, 2, 3, 5, {0, 0, 0, 128, 136}},
    {128, 16, 4, 0, 1, {3, 15, 1, 143, 148}},   {128, 16, 4, 0, 4, {3, 15, 1, 143, 148}},
    {128, 16, 4, 1, 4, {3, 15, 1, 143, 148}},   {128, 8, 8, 0, 3, {4, 7, 1, 135, 138}},
    {256, 64, 1, 0, 2, {2, 63, 1, 319, 319}},   {256, 64, 1, 0, 6, {4, 15, 1, 271, 271}},
    {256, 64, 1, 2, 4, {4, 15, 4, 316, 316}},   {256, 64, 1, 4, 6, {0, 0, 0, 256, 256}},
    {256, 32, 2, 0, 2, {3, 31, 1, 287, 296}},   {256, 32, 2, 0, 5, {3, 31, 1, 287, 296}},
    {256, 32, 2, 2, 5, {3, 31, 1, 287, 296}},   {256, 16, 4, 0, 4, {4, 15, 1, 271, 276}},
    {256, 8, 8, 0, 3, {5, 7, 1, 263, 266}},     {512, 64, 1, 0, 3, {3, 63, 1, 575, 575}},
    {512, 64, 1, 0, 6, {4, 31, 1, 543, 543}},   {512, 64, 1, 3, 6, {4, 31, 2, 574, 574}},
    {512, 32, 2, 0, 1, {4, 31, 1, 543, 552}},   {512, 32, 2, 0, 5, {4, 31, 1, 543, 552}},
    {512, 32, 2, 1, 5, {4, 31, 1, 543, 552}},   {512, 16, 4, 0, 4, {5, 15, 1, 527, 532}},
    {512, 8, 8, 0, 3, {6, 7, 1, 519, 522}},
};
constexpr TPad tpad(int N, int P, int LS, int k1, int k2) {
  const int a = k1 < k2 ? k1 : k2, b = k1 < k2 ? k2 : k1;
  for (const TPadEntry& e : kTPad)
    if (e.N == N && e.P == P && e.LS == LS && e.a == a && e.b == b) return e.pad;
  return {-1, 0, 0, 0, 0};
}

template <int N, int P, int LS>
struct TPlan {
  static constexpr int L = N / P, LG = ilog2c(L), M = ilog2c(N), K0 = M - LG;
  // layout j of the forward sequence K0, K0 - LG, ... (clamped at 0)
  static constexpr int kseq(int j) {
    int k = K0;
    for (int i = 0; i < j; ++i) k = k - LG > 0 ? k - LG : 0;
    return k;
  }
  static constexpr int nlayouts() {
    if (LG == 0) return 1;  // one value per lane: the cross-lane form
    int j = 0;
    while (kseq(j) > 0) ++j;
    return j + 1;
  }
  static constexpr int NK = nlayouts();
  static constexpr bool ENABLED = L >= 4;
  // scratch elements per wave: LS lines of the largest stride used
  static constexpr int scratch() {
    int m = 0;
    for (int j = 0; j < NK; ++j) {
      const TPad a = tpad(N, P, LS, kseq(j), kseq(j + 1 < NK ? j + 1 : 0));
      const int e = (LS == 1 ? a.size : a.stride * LS);
      m = e > m ? e : m;
    }
    return m;
  }
  static constexpr int SCRATCH = ENABLED ? scratch() : 0;
};

__host__ __device__ constexpr int tpos(int l, int t, int k, int lg) {
  return ((t >> k) << (k + lg)) | (l << k) | (t & ((1 << k) - 1));
}
__host__ __device__ constexpr int tpadded(const TPad& d, int n) { return n + d.cc * ((n >> d.s) & d.msk); }

// one-to-one and conflict-free for both layouts (compile-time check)
constexpr bool tpad_ok(int N, int P, int LS, int k1, int k2) {
  const TPad d = tpad(N, P, LS, k1, k2);
  if (d.s < 0) return false;
  const int lg = ilog2c(N / P);
  // one-to-one into [0, size): every padded index below size and distinct
  if (d.size > 2048) return false;
  bool seen[2048] = {};
  for (int n = 0; n < N; ++n) {
    const int a = tpadded(d, n);
    if (a < 0 || a >= d.size || seen[a]) return false;
    seen[a] = true;
  }
  if (LS > 1 && d.stride < d.size) return false;
  const int ks[2] = {k1, k2};
  for (int ki = 0; ki < 2; ++ki)
    for (int l = 0; l < N / P; ++l)
      for (int g = 0; g < 4; ++g) {
        unsigned used = 0;
        for (int lane = 16 * g; lane < 16 * g + 16; ++lane) {
          const int t = lane / LS, c = lane % LS;
          const int bank = (c * d.stride + tpadded(d, tpos(l, t, ks[ki], lg))) & 15;
          if (used & (1u << bank)) return false;
          used |= 1u << bank;
        }
      }
  return true;
}

template <int N, int P, int LS>
constexpr bool tplan_ok() {
  using T = TPlan<N, P, LS>;
  if (!T::ENABLED) return true;
  for (int j = 0; j < T::NK; ++j)
    if (!tpad_ok(N, P, LS, T::kseq(j), T::kseq(j + 1 < T::NK ? j + 1 : 0))) return false;
  return true;
}

// Lane part of layout k's element index: n(l, t; k) = tpart + (l << k).
template <int LG>
__device__ __forceinline__ int tpart(int t, int k) {
  return ((t >> k) << (k + LG)) | (t & ((1 << k) - 1));
}

// Rewrite the lane's L values from layout K1 to layout K2 through the
// wave-private scratch `sx` (line c of the wave at c * stride).  LDS
// instructions of one wave are performed in order, so the reads see the
// other lanes' writes; the compiler barriers keep the accesses in program
// order.
// Swap of lane bit 5 (m = 32) or lane bit 4 (m = 16) with the register bit
// that tells a from b: the gfx950 permlane swaps exchange the upper half of a
// with the lower half of b (m = 32), or the odd rows of 16 lanes of a with the
// even rows of b (m = 16), so afterwards a holds the values whose lane bit was
// 0, with that lane bit now carrying the register bit, and b those whose lane
// bit was 1.  Exact data movement.
template <int M>
__device__ __forceinline__ void lane_reg_swap(float2& a, float2& b) {
  static_assert(M == 16 || M == 32, "lane bit 4 or 5");
  const int ax = __float_as_int(a.x), ay = __float_as_int(a.y), bx = __float_as_int(b.x), by = __float_as_int(b.y);
  if constexpr (M == 32) {
    const auto px = __builtin_amdgcn_permlane32_swap(ax, bx, false, false);
    const auto py = __builtin_amdgcn_permlane32_swap(ay, by, false, false);
    a = make_float2(__int_as_float((int)px[0]), __int_as_float((int)py[0]));
    b = make_float2(__int_as_float((int)px[1]), __int_as_float((int)py[1]));
  } else {
    const auto px = __builtin_amdgcn_permlane16_swap(ax, bx, false, false);
    const auto py = __builtin_amdgcn_permlane16_swap(ay, by, false, false);
    a = make_float2(__int_as_float((int)px[0]), __int_as_float((int)py[0]));
    b = make_float2(__int_as_float((int)px[1]), __int_as_float((int)py[1]));
  }
}

template <int N, int P, int LS, int K1, int K2>
__device__ __forceinline__ void ttranspose(float2 (&v)[N / P], float2* __restrict__ sx, int t, int c) {
  using T = TPlan<N, P, LS>;
  if constexpr (N == 256 && P == 64 && LS == 1 && ((K1 == 6 && K2 == 4) || (K1 == 4 && K2 == 6))) {
    // Layouts 6 and 4 of a 256-point line on 64 lanes differ only in which of
    // the element index bits 4, 5 and 6, 7 sit in the lane (bits 4, 5) and
    // which in the register index (bits 0, 1): swap lane bit 5 with register
    // bit 1 and lane bit 4 with register bit 0, in registers (no LDS).
    lane_reg_swap<32>(v[0], v[2]);
    lane_reg_swap<32>(v[1], v[3]);
    lane_reg_swap<16>(v[0], v[1]);
    lane_reg_swap<16>(v[2], v[3]);
    (void)sx;
    (void)t;
    (void)c;
    return;
  }
  constexpr TPad d = tpad(N, P, LS, K1, K2);
  static_assert(d.s >= 0, "no transpose padding for this layout pair");
  float2* base = sx + c * d.stride;
  const int t1 = tpart<T::LG>(t, K1), t2 = tpart<T::LG>(t, K2);
  float2* w = base + tpadded(d, t1);
  const float2* r = base + tpadded(d, t2);
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int l = 0; l < T::L; ++l) w[tpadded(d, l << K1)] = v[l];
  __asm__ volatile("" ::: "memory");
#pragma unroll
  for (int l = 0; l < T::L; ++l) v[l] = r[tpadded(d, l << K2)];
  __asm__ volatile("" ::: "memory");
}

// Forward (DIF) stages of bits hi-1 ... lo inside layout K.
template <int N, int P, int LS, int K>
__device__ __forceinline__ void tstages_fwd(float2 (&v)[N / P], const float2* __restrict__ tw, int t, int hi, int lo) {
  using T = TPlan<N, P, LS>;
  const int tlo = t & ((1 << K) - 1);
#pragma unroll
  for (int b = K + T::LG - 1; b >= K; --b) {
    if (b >= hi || b < lo) continue;
    const int sh = T::M - 1 - b;
    const float2* __restrict__ twb = tw + (tlo << sh);
    const int d = 1 << (b - K);
#pragma unroll
    for (int l = 0; l < T::L; ++l) {
      if ((l & d) == 0) {
        const float2 w = twb[((l & (d - 1)) << K) << sh];
        const float2 a = v[l], q = v[l + d];
        v[l] = cadd(a, q);
        v[l + d] = cmul_f(csub(a, q), w);
      }
    }
  }
}

// Inverse (DIT) stages of bits lo ... hi-1 inside layout K.
template <int N, int P, int LS, int K>
__device__ __forceinline__ void tstages_inv(float2 (&v)[N / P], const float2* __restrict__ tw, int t, int hi, int lo) {
  using T = TPlan<N, P, LS>;
  const int tlo = t & ((1 << K) - 1);
#pragma unroll
  for (int b = K; b < K + T::LG; ++b) {
    if (b >= hi || b < lo) continue;
    const int sh = T::M - 1 - b;
    const float2* __restrict__ twb = tw + (tlo << sh);
    const int d = 1 << (b - K);
#pragma unroll
    for (int l = 0; l < T::L; ++l) {
      if ((l & d) == 0) {
        const float2 w = conjf2(twb[((l & (d - 1)) << K) << sh]);
        const float2 a = v[l];
        const float2 q = cmul_f(v[l + d], w);
        v[l] = cadd(a, q);
        v[l + d] = csub(a, q);
      }
    }
  }
}

template <int N, int P, int LS, int J, bool Z>
__device__ __forceinline__ void tfwd_from(float2 (&v)[N / P], const float2* __restrict__ tw, float2* __restrict__ sx,
                                          int t, int c) {
  using T = TPlan<N, P, LS>;
  constexpr int K = T::kseq(J);
  constexpr int HI = J == 0 ? T::M : T::kseq(J - 1);
  tstages_fwd<N, P, LS, K>(v, tw, t, HI, K);
  if constexpr (J + 1 < T::NK) {
    ttranspose<N, P, LS, K, T::kseq(J + 1)>(v, sx, t, c);
    tfwd_from<N, P, LS, J + 1, Z>(v, tw, sx, t, c);
  } else if constexpr (!Z) {
    ttranspose<N, P, LS, K, T::K0>(v, sx, t, c);
  }
}

template <int N, int P, int LS, int J>
__device__ __forceinline__ void tinv_from(float2 (&v)[N / P], const float2* __restrict__ tw, float2* __restrict__ sx,
                                          int t, int c) {
  // layouts in reverse order: J = NK - 1 first (lowest bits), J = 0 last
  using T = TPlan<N, P, LS>;
  constexpr int K = T::kseq(J);
  constexpr int HI = J == 0 ? T::M : T::kseq(J - 1);
  tstages_inv<N, P, LS, K>(v, tw, t, HI, K);
  if constexpr (J > 0) {
    ttranspose<N, P, LS, K, T::kseq(J - 1)>(v, sx, t, c);
    tinv_from<N, P, LS, J - 1>(v, tw, sx, t, c);
  }
}

// fft_fwd / fft_inv for a line on P lanes (lane stride LS): the transposed
// form where it is enabled (L >= 4 values per lane; `sx` then points to the
// calling wave's TPlan::SCRATCH elements of LDS, `c` = lane % LS), the
// cross-lane form otherwise.  With Z the frequency side (the forward
// output, the inverse input) stays in the transposed form's last layout,
// where register l of lane t holds slot zslot(l, t) of the line, and the
// transpose between that layout and the caller's is skipped; without Z
// both sides use the caller's layout, slot l * P + t.
template <int N, int P, int LS>
__device__ __forceinline__ int zslot(int l, int t) {
  using T = TPlan<N, P, LS>;
  if constexpr (T::ENABLED) {
    static_assert(T::kseq(T::NK - 1) == 0, "the last layout of the transposed form is layout 0");
    return (t << T::LG) | l;
  } else {
    return l * P + t;
  }
}
template <int N, int LPF, int LS, bool Z = false>
__device__ __forceinline__ void fft_fwd_g(float2 (&v)[N / LPF], const float2* __restrict__ tw_, int t,
                                          float2* __restrict__ sx, int c) {
  if constexpr (TPlan<N, LPF, LS>::ENABLED) {
    static_assert(tplan_ok<N, LPF, LS>(), "transpose paddings are not conflict-free");
    unsigned off = 0;
    __asm__ volatile("" : "+s"(off));
    tfwd_from<N, LPF, LS, 0, Z>(v, tw_ + off, sx, t, c);
  } else {
    fft_fwd_x<N, LPF, LS>(v, tw_, t);
  }
}
template <int N, int LPF, int LS, bool Z = false>
__device__ __forceinline__ void fft_inv_g(float2 (&v)[N / LPF], const float2* __restrict__ tw_, int t,
                                          float2* __restrict__ sx, int c) {
  if constexpr (TPlan<N, LPF, LS>::ENABLED) {
    static_assert(tplan_ok<N, LPF, LS>(), "transpose paddings are not conflict-free");
    using T = TPlan<N, LPF, LS>;
    unsigned off = 0;
    __asm__ volatile("" : "+s"(off));
    // the inverse starts in the last layout
    if constexpr (!Z) ttranspose<N, LPF, LS, T::K0, T::kseq(T::NK - 1)>(v, sx, t, c);
    tinv_from<N, LPF, LS, T::NK - 1>(v, tw_ + off, sx, t, c);
  } else {
    fft_inv_x<N, LPF, LS>(v, tw_, t);
  }
}

// per-wave scratch (float2 elements) the transforms of a probe size need:
// rows (LPF lanes) and column groups of CW lines
template <int N, int CW>
constexpr int fft_scratch() {
  constexpr int LPF = N >= 64 ? 64 : N;
  constexpr int cw = CW * N >= 64 ? CW : 64 / N;
  constexpr int a = TPlan<N, LPF, 1>::SCRATCH;
  constexpr int b = TPlan<N, 64 / cw, cw>::SCRATCH;
  constexpr int m = a > b ? a : b;
  return m > 0 ? m : 1;
}

// row layout: a line on LPF adjacent lanes (Z: frequency side in slots
// rslot(l, t), see zslot)
template <int N, bool Z = false>
__device__ __forceinline__ void fft_fwd(float2 (&v)[Line<N>::L], const float2* __restrict__ tw, int t, float2* sx) {
  fft_fwd_g<N, Line<N>::LPF, 1, Z>(v, tw, t, sx, 0);
}
template <int N, bool Z = false>
__device__ __forceinline__ void fft_inv(float2 (&v)[Line<N>::L], const float2* __restrict__ tw, int t, float2* sx) {
  fft_inv_g<N, Line<N>::LPF, 1, Z>(v, tw, t, sx, 0);
}
template <int N>
__device__ __forceinline__ int rslot(int l, int t) { return zslot<N, Line<N>::LPF, 1>(l, t); }

__device__ __forceinline__ int brev(int x, int logn) { return (int)(__brev((unsigned)x) >> (32 - logn)); }

}  // namespace pty
