// Wavefront-resident FFT building blocks for gfx950 (wave64).
//
// A length-N line lives in registers of a group of LPF lanes: element
// n = l * LPF + t sits in register l of the lane whose in-line index is t
// (L = N / LPF registers per lane).  Consecutive in-line indices are LS
// lanes apart in the wave (LS = 1: a line occupies adjacent lanes; LS = CW:
// CW lines interleaved lane by lane, as the column passes use).
// Butterflies whose span is a multiple of LPF are done inside a lane; smaller
// spans exchange values between lanes (mask span * LS) with DPP lane moves
// and the permlane swaps (xlane).  No LDS and no workgroup barrier is needed
// inside a transform.
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
__device__ __forceinline__ void fft_fwd_g(float2 (&v)[N / LPF], const float2* __restrict__ tw_, int t) {
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
__device__ __forceinline__ void fft_inv_g(float2 (&v)[N / LPF], const float2* __restrict__ tw_, int t) {
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

// row layout: a line on LPF adjacent lanes
template <int N>
__device__ __forceinline__ void fft_fwd(float2 (&v)[Line<N>::L], const float2* __restrict__ tw, int t) {
  fft_fwd_g<N, Line<N>::LPF, 1>(v, tw, t);
}
template <int N>
__device__ __forceinline__ void fft_inv(float2 (&v)[Line<N>::L], const float2* __restrict__ tw, int t) {
  fft_inv_g<N, Line<N>::LPF, 1>(v, tw, t);
}

__device__ __forceinline__ int brev(int x, int logn) { return (int)(__brev((unsigned)x) >> (32 - logn)); }

}  // namespace pty
