// Learned initialiser kernels for gfx950: the layers of a small U-Net
// (3x3 / 1x1 convolutions, 2x2 average pooling, nearest 2x upsampling,
// channel concatenation), their backward passes, an MSE loss and Adam.
//
// Tensors are NHWC with a channel stride that is a multiple of 8 (16-byte
// rows of bf16); padding channels hold zeros and their weights stay zero.
// Activations and gradients are bf16, weights have an fp32 master copy and
// a bf16 copy for the matrix cores; every product accumulates in fp32 on
// the MFMA units (v_mfma_f32_32x32x16_bf16).
//
// Convolutions are implicit GEMMs:
//   forward      Y[p, co]  = sum_k  im2col(X)[p, k] * W[co, k]       (M = pixels)
//   input grad   dX        = forward convolution of dZ with the flipped,
//                            transposed weights W'[ci, (ky,kx,co)]
//   weight grad  dW[co, k] = sum_p  dZ[p, co] * im2col(X)[p, k]     (split over p)
// with k = (ky * ks + kx) * cin_s + ci.  The bias gradient rides along the
// weight gradient as an extra column k = K whose im2col value is 1.
// Weight rows are padded to Kp = round_up(K + 1, 32) with zeros.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>

#include "pty_net.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint16_t bf16_t;  // storage type: raw bfloat16 bits

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

#define HIP_TRY(expr)                                                             \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return set_err(PTY_NET_EDEVICE, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

}  // namespace

// float -> bfloat16, round to nearest even (NaN kept quiet)
__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (bf16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ---------------------------------------------------------------------------
// MFMA operand layout of v_mfma_f32_32x32x16_bf16 (one wave, 64 lanes):
//   A (32 x 16): lane l holds A[l % 32][8 * (l / 32) + v], v = 0..7
//   B (16 x 32): lane l holds B[8 * (l / 32) + v][l % 32]
//   D (32 x 32): lane l, register r holds D[8 * (r / 4) + 4 * (l / 32) + r % 4][l % 32]
// pty_net_mfma_probe runs one instruction on caller-chosen operands so the
// layout is checked against a host product.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int d_row(int lane, int r) { return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

__global__ __launch_bounds__(64) void mfma_probe_kernel(const bf16_t* __restrict__ A, const bf16_t* __restrict__ B,
                                                        float* __restrict__ D) {
  const int l = threadIdx.x;
  bf16x8 a, b;
  for (int v = 0; v < 8; ++v) {
    a[v] = __builtin_bit_cast(__bf16, A[(l & 31) * 16 + 8 * (l >> 5) + v]);
    b[v] = __builtin_bit_cast(__bf16, B[(8 * (l >> 5) + v) * 32 + (l & 31)]);
  }
  f32x16 acc = {};
  acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 16; ++r) D[d_row(l, r) * 32 + (l & 31)] = acc[r];
}

// ---------------------------------------------------------------------------
// implicit-GEMM convolution, forward (also used for the input gradient)
// ---------------------------------------------------------------------------
#define CV_BM 128  // pixels per workgroup tile
#define CV_BN 64   // output channels per workgroup tile
#define CV_KT 32   // reduction step
#define CV_LD 40   // LDS row stride (bf16): 32 + 8 keeps 16-byte reads off one bank group

struct ConvFwd {
  const bf16_t* x;     // (B, H, W, cin_s)
  const bf16_t* w;     // (cout, Kp)
  const float* bias;   // (cout) or null
  const bf16_t* gate;  // (B, H, W, cout_s) or null: y *= (gate > 0)
  void* y;             // (B, H, W, cout_s), bf16 or f32
  int B, H, W, cin_s, cout, cout_s, ks, K, Kp, relu, out_f32;
};

__device__ __forceinline__ uint4 im2col_chunk(const bf16_t* __restrict__ x, int m, int k, int M, int H, int W, int cin_s,
                                              int ks, int K, bool ones_at_K) {
  uint4 z = make_uint4(0u, 0u, 0u, 0u);
  if (m >= M) return z;
  if (k >= K) {
    // the bias column of the weight gradient: value 1 at k == K
    if (ones_at_K && k <= K && K < k + 8) {
      const int pos = K - k;
      uint32_t words[4] = {0u, 0u, 0u, 0u};
      words[pos >> 1] = (pos & 1) ? 0x3f800000u : 0x00003f80u;
      return make_uint4(words[0], words[1], words[2], words[3]);
    }
    return z;
  }
  const int tap = k / cin_s, ci = k - tap * cin_s;
  const int ky = tap / ks, kx = tap - ky * ks;
  const int hw = H * W;
  const int n = m / hw, r = m - n * hw;
  const int y = r / W, xx = r - y * W;
  const int yy = y + ky - ks / 2, xs = xx + kx - ks / 2;
  if (yy < 0 || yy >= H || xs < 0 || xs >= W) return z;
  return *(const uint4*)(x + ((size_t)(n * H + yy) * W + xs) * cin_s + ci);
}

__global__ __launch_bounds__(256) void conv_fwd_kernel(ConvFwd a) {
  __shared__ __attribute__((aligned(16))) bf16_t As[CV_BM * CV_LD];
  __shared__ __attribute__((aligned(16))) bf16_t Bs[CV_BN * CV_LD];
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int M = a.B * a.H * a.W;
  const int m0 = blockIdx.x * CV_BM, n0 = blockIdx.y * CV_BN;
  // This thread fills A-tile rows t / 4 and t / 4 + 64 (two pixels) at k
  // chunk t % 4 of every 32-wide step: the pixels' coordinates are computed
  // once and the chunk's (tap, channel) advances by 32 per step, so the
  // loop does no integer division (cin_s is a multiple of 8: a chunk never
  // straddles two taps).  Same operands as im2col_chunk.
  const int kc = t & 3, row0 = t >> 2;
  const int hw = a.H * a.W, half = a.ks / 2;
  const bf16_t* img[2];
  int py[2], pxx[2];
  bool pin[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int m = m0 + row0 + 64 * h;
    pin[h] = m < M;
    const int mm = pin[h] ? m : 0;
    const int n = mm / hw, r = mm - n * hw;
    py[h] = r / a.W;
    pxx[h] = r - py[h] * a.W;
    img[h] = a.x + (size_t)n * hw * a.cin_s;
  }
  int tap = (8 * kc) / a.cin_s;
  int ci = 8 * kc - tap * a.cin_s;
  int ky = tap / a.ks, kx = tap - ky * a.ks;
  f32x16 acc0 = {}, acc1 = {};
  // Software pipeline: the chunks of step k0 + 32 are loaded into registers
  // while the MFMAs work on step k0 in LDS.
  uint4 ra[2], rb;
  const int brow = t >> 2, co_b = n0 + brow;
  auto fetch = [&](int kb) {
    const bool kin = kb + 8 * kc < a.K;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      ra[h] = make_uint4(0u, 0u, 0u, 0u);
      const int yy = py[h] + ky - half, xs = pxx[h] + kx - half;
      if (pin[h] && kin && yy >= 0 && yy < a.H && xs >= 0 && xs < a.W)
        ra[h] = *(const uint4*)(img[h] + ((size_t)yy * a.W + xs) * a.cin_s + ci);
    }
    ci += CV_KT;
    while (ci >= a.cin_s) {
      ci -= a.cin_s;
      if (++kx == a.ks) {
        kx = 0;
        ++ky;
      }
    }
    rb = make_uint4(0u, 0u, 0u, 0u);
    if (co_b < a.cout) rb = *(const uint4*)(a.w + (size_t)co_b * a.Kp + kb + 8 * kc);
  };
  fetch(0);
  for (int k0 = 0; k0 < a.Kp; k0 += CV_KT) {
    // A tile: 128 pixels x 32 k, two 8-element chunks per thread; B tile:
    // 64 output channels x 32 k, one chunk per thread
#pragma unroll
    for (int h = 0; h < 2; ++h) *(uint4*)(As + (row0 + 64 * h) * CV_LD + 8 * kc) = ra[h];
    *(uint4*)(Bs + brow * CV_LD + 8 * kc) = rb;
    __syncthreads();
    if (k0 + CV_KT < a.Kp) fetch(k0 + CV_KT);
#pragma unroll
    for (int kk = 0; kk < CV_KT; kk += 16) {
      const bf16x8 av = *(const bf16x8*)(As + (32 * wave + (l & 31)) * CV_LD + kk + 8 * (l >> 5));
      const bf16x8 b0 = *(const bf16x8*)(Bs + (l & 31) * CV_LD + kk + 8 * (l >> 5));
      const bf16x8 b1 = *(const bf16x8*)(Bs + (32 + (l & 31)) * CV_LD + kk + 8 * (l >> 5));
      acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b0, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, b1, acc1, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int nt = 0; nt < 2; ++nt) {
    const int co = n0 + 32 * nt + (l & 31);
    if (co >= a.cout) continue;
    const float bb = a.bias ? a.bias[co] : 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int m = m0 + 32 * wave + d_row(l, r);
      if (m >= M) continue;
      float v = (nt == 0 ? acc0[r] : acc1[r]) + bb;
      if (a.relu) v = v > 0.f ? v : 0.f;
      const size_t o = (size_t)m * a.cout_s + co;
      if (a.gate && !(bf2f(a.gate[o]) > 0.f)) v = 0.f;
      if (a.out_f32)
        ((float*)a.y)[o] = v;
      else
        ((bf16_t*)a.y)[o] = f2bf(v);
    }
  }
}

// ---------------------------------------------------------------------------
// weight gradient: dW[co, k] = sum_p dZ[p, co] * im2col(X)[p, k], split over
// pixel ranges into fp32 partials, reduced in a fixed order
// ---------------------------------------------------------------------------
#define WG_BM 64  // output channels per tile
#define WG_BN 64  // k per tile
#define WG_PT 32  // pixels per step

struct ConvWgrad {
  const bf16_t* x;   // (B, H, W, cin_s) layer input
  const bf16_t* dz;  // (B, H, W, cout_s) gradient at the pre-activation
  float* part;       // (splits, cout, Kp)
  int B, H, W, cin_s, cout, cout_s, ks, K, Kp, pix_per_split;
};

__global__ __launch_bounds__(256) void conv_wgrad_kernel(ConvWgrad a) {
  __shared__ __attribute__((aligned(16))) bf16_t At[WG_BM * CV_LD];  // [co][p]
  __shared__ __attribute__((aligned(16))) bf16_t Bt[WG_BN * CV_LD];  // [k][p]
  const int t = threadIdx.x, wave = t >> 6, l = t & 63;
  const int wm = wave >> 1, wn = wave & 1;
  const int M = a.B * a.H * a.W;
  const int k0 = blockIdx.x * WG_BN, co0 = blockIdx.y * WG_BM;
  const int pb = blockIdx.z * a.pix_per_split;
  const int pe = pb + a.pix_per_split < M ? pb + a.pix_per_split : M;
  f32x16 acc = {};
  for (int p0 = pb; p0 < pe; p0 += WG_PT) {
    {
      // dZ tile: 32 pixels x 64 channels, one 8-channel chunk per thread
      const int p = t >> 3, cc = t & 7, m = p0 + p, co = co0 + 8 * cc;
      uint4 v = make_uint4(0u, 0u, 0u, 0u);
      if (m < pe && co < a.cout) v = *(const uint4*)(a.dz + (size_t)m * a.cout_s + co);
      const bf16_t* e = (const bf16_t*)&v;
#pragma unroll
      for (int q = 0; q < 8; ++q) At[(8 * cc + q) * CV_LD + p] = e[q];
    }
    {
      // im2col tile: 32 pixels x 64 k
      const int p = t >> 3, kc = t & 7, m = p0 + p;
      const uint4 v = im2col_chunk(a.x, m < pe ? m : M, k0 + 8 * kc, M, a.H, a.W, a.cin_s, a.ks, a.K, true);
      const bf16_t* e = (const bf16_t*)&v;
#pragma unroll
      for (int q = 0; q < 8; ++q) Bt[(8 * kc + q) * CV_LD + p] = e[q];
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < WG_PT; kk += 16) {
      const bf16x8 av = *(const bf16x8*)(At + (32 * wm + (l & 31)) * CV_LD + kk + 8 * (l >> 5));
      const bf16x8 bv = *(const bf16x8*)(Bt + (32 * wn + (l & 31)) * CV_LD + kk + 8 * (l >> 5));
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int k = k0 + 32 * wn + (l & 31);
  float* out = a.part + (size_t)blockIdx.z * a.cout * a.Kp;
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int co = co0 + 32 * wm + d_row(l, r);
    if (co < a.cout && k < a.Kp) out[(size_t)co * a.Kp + k] = acc[r];
  }
}

__global__ void wgrad_reduce_kernel(const float* __restrict__ part, int splits, int cout, int Kp, int K,
                                    float* __restrict__ dw, float* __restrict__ db) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cout * Kp) return;
  float s = 0.f;
  for (int q = 0; q < splits; ++q) s += part[(size_t)q * cout * Kp + i];
  const int co = i / Kp, k = i - co * Kp;
  if (k == K) {
    if (db) db[co] = s;
    dw[i] = 0.f;
  } else {
    dw[i] = k < K ? s : 0.f;
  }
}

// W'[ci, (ky', kx', co)] = W[co, (ks-1-ky', ks-1-kx', ci)]: the weights of the
// input-gradient convolution (input channel stride cout_s, K' = ks^2 cout_s)
__global__ void weight_flip_kernel(const bf16_t* __restrict__ w, int Kp, int cin, int cin_s, int cout, int cout_s, int ks,
                                   bf16_t* __restrict__ wt, int Kpt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cin * Kpt) return;
  const int ci = i / Kpt, k = i - ci * Kpt;
  bf16_t v = 0;
  if (k < ks * ks * cout_s) {
    const int tap = k / cout_s, co = k - tap * cout_s;
    const int ky = tap / ks, kx = tap - ky * ks;
    if (co < cout) v = w[(size_t)co * Kp + ((ks - 1 - ky) * ks + (ks - 1 - kx)) * cin_s + ci];
  }
  wt[i] = v;
}

// ---------------------------------------------------------------------------
// elementwise layers (NHWC, channel stride c_s)
// ---------------------------------------------------------------------------
__global__ void pool_fwd_kernel(const bf16_t* __restrict__ x, bf16_t* __restrict__ y, int B, int H, int W, int c_s) {
  const int h = H / 2, w = W / 2;
  const size_t n = (size_t)B * h * w * c_s;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % c_s);
    size_t r = i / c_s;
    const int xo = (int)(r % w);
    r /= w;
    const int yo = (int)(r % h);
    const int b = (int)(r / h);
    const size_t base = (((size_t)b * H + 2 * yo) * W + 2 * xo) * c_s + c;
    const float s = (bf2f(x[base]) + bf2f(x[base + c_s])) + (bf2f(x[base + (size_t)W * c_s]) + bf2f(x[base + (size_t)W * c_s + c_s]));
    y[i] = f2bf(0.25f * s);
  }
}

__global__ void pool_bwd_kernel(const bf16_t* __restrict__ dy, bf16_t* __restrict__ dx, int B, int H, int W, int c_s) {
  const size_t n = (size_t)B * H * W * c_s;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % c_s);
    size_t r = i / c_s;
    const int x = (int)(r % W);
    r /= W;
    const int y = (int)(r % H);
    const int b = (int)(r / H);
    dx[i] = f2bf(0.25f * bf2f(dy[(((size_t)b * (H / 2) + y / 2) * (W / 2) + x / 2) * c_s + c]));
  }
}

// nearest 2x upsampling of x (B, h, w, ca_s) into channels [0, ca_s) of the
// concatenation y (B, 2h, 2w, ca_s + cb_s); channels [ca_s, ...) copy skip
__global__ void up_concat_kernel(const bf16_t* __restrict__ xa, int ca_s, const bf16_t* __restrict__ xb, int cb_s,
                                 bf16_t* __restrict__ y, int B, int H, int W) {
  // one thread moves 8 channels (16 bytes); channel strides are multiples of
  // 8 and the host keeps the chunk count below 2^31 (32-bit index math)
  const unsigned ca8 = (unsigned)ca_s / 8u, cs8 = (unsigned)(ca_s + cb_s) / 8u;
  const unsigned n = (unsigned)(B * H * W) * cs8;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const unsigned c8 = i % cs8, pix = i / cs8;
    uint4 v;
    if (c8 < ca8) {
      const unsigned x = pix % (unsigned)W, r = pix / (unsigned)W;
      const unsigned yy = r % (unsigned)H, b = r / (unsigned)H;
      v = *(const uint4*)(xa + (((size_t)b * (H / 2) + yy / 2) * (W / 2) + x / 2) * ca_s + 8 * c8);
    } else {
      v = *(const uint4*)(xb + (size_t)pix * cb_s + 8 * (c8 - ca8));
    }
    *(uint4*)(y + (size_t)i * 8) = v;
  }
}

// backward of up_concat: dxa (B, h, w, ca_s) = sum of the four upsampled
// gradients; dxb (B, H, W, cb_s) += the skip part (accumulated in fp32 when
// the skip tensor already has a gradient from its own consumer)
__global__ void up_concat_bwd_kernel(const bf16_t* __restrict__ dy, int ca_s, int cb_s, bf16_t* __restrict__ dxa,
                                     const bf16_t* __restrict__ dxb_in, bf16_t* __restrict__ dxb, int B, int H, int W) {
  const int cs = ca_s + cb_s;
  const int h = H / 2, w = W / 2;
  const size_t na = (size_t)B * h * w * ca_s;
  const size_t nb = (size_t)B * H * W * cb_s;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < na + nb; i += (size_t)gridDim.x * blockDim.x) {
    if (i < na) {
      const int c = (int)(i % ca_s);
      size_t r = i / ca_s;
      const int x = (int)(r % w);
      r /= w;
      const int yy = (int)(r % h);
      const int b = (int)(r / h);
      const size_t p = ((size_t)b * H + 2 * yy) * W + 2 * x;
      const float s = (bf2f(dy[p * cs + c]) + bf2f(dy[(p + 1) * cs + c])) +
                      (bf2f(dy[(p + W) * cs + c]) + bf2f(dy[(p + W + 1) * cs + c]));
      dxa[i] = f2bf(s);
    } else {
      const size_t j = i - na;
      const int c = (int)(j % cb_s);
      const size_t pix = j / cb_s;
      float v = bf2f(dy[pix * cs + ca_s + c]);
      if (dxb_in) v += bf2f(dxb_in[j]);
      dxb[j] = f2bf(v);
    }
  }
}

// y = relu'(a) * (g1 + g2): combines two gradient contributions at an
// activation a = relu(z) and applies the ReLU derivative
__global__ void add_gate_kernel(const bf16_t* __restrict__ g1, const bf16_t* __restrict__ g2,
                                const bf16_t* __restrict__ act, bf16_t* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float v = bf2f(g1[i]) + (g2 ? bf2f(g2[i]) : 0.f);
    y[i] = (act && !(bf2f(act[i]) > 0.f)) ? (bf16_t)0 : f2bf(v);
  }
}

// MSE over the crop [oy, oy + hc) x [ox, ox + wc) of (B, Ho, Wo, 2) fp32
// outputs against tgt (B, hc, wc, 2); dout (B, Ho, Wo, c_s) bf16 gets
// 2 (out - t) / n inside the crop (n = B hc wc 2; outside stays as it is,
// zero); per-block FP64 partial sums of the squared error
__global__ void mse_kernel(const float* __restrict__ out, const float* __restrict__ tgt, int B, int Ho, int Wo, int oy,
                           int ox, int hc, int wc, int c_s, bf16_t* __restrict__ dout, double* __restrict__ partial) {
  __shared__ double red[256];
  const size_t n = (size_t)B * hc * wc * 2;
  const float scale = 2.0f / (float)n;
  double acc = 0.0;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const int c = (int)(i & 1);
    size_t r = i >> 1;
    const int x = (int)(r % wc);
    r /= wc;
    const int y = (int)(r % hc);
    const int b = (int)(r / hc);
    const size_t pix = ((size_t)b * Ho + oy + y) * Wo + ox + x;
    const float d = out[pix * 2 + c] - tgt[i];
    acc += (double)d * (double)d;
    dout[pix * c_s + c] = f2bf(scale * d);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (Kingma & Ba): fp32 master weights, bf16 copy for the matrix cores
__global__ void adam_kernel(float* __restrict__ w, bf16_t* __restrict__ w16, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, size_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = f2bf(wi);
  }
}

__global__ void cast_bf16_kernel(const float* __restrict__ x, bf16_t* __restrict__ y, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {
int grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 16384) g = 16384;
  return (int)(g ? g : 1);
}
int check_conv_shape(int B, int H, int W, int cin_s, int cout, int cout_s, int ks, int Kp) {
  if (B < 1 || H < 1 || W < 1 || cout < 1) return set_err(PTY_NET_EINVAL, "conv: empty shape");
  if (cin_s % 8 || cout_s % 8 || cin_s < 8) return set_err(PTY_NET_EINVAL, "conv: channel strides must be multiples of 8");
  if (cout > cout_s) return set_err(PTY_NET_EINVAL, "conv: cout %d > stride %d", cout, cout_s);
  if (ks != 1 && ks != 3) return set_err(PTY_NET_EINVAL, "conv: kernel size %d (1 or 3)", ks);
  if (Kp != pty_net_kpad(ks, cin_s)) return set_err(PTY_NET_EINVAL, "conv: weight row stride %d != %d", Kp, pty_net_kpad(ks, cin_s));
  if ((size_t)B * H * W >= (size_t)1 << 31) return set_err(PTY_NET_EINVAL, "conv: too many pixels");
  return PTY_NET_OK;
}
}  // namespace

extern "C" {

const char* pty_net_last_error(void) { return g_err.c_str(); }

int pty_net_kpad(int ks, int cin_s) {
  const int K = ks * ks * cin_s;
  return (K + 1 + 31) / 32 * 32;
}

int pty_net_mfma_probe(const void* A, const void* B, float* D, void* stream) {
  if (!A || !B || !D) return set_err(PTY_NET_EINVAL, "mfma_probe: null pointer");
  mfma_probe_kernel<<<1, 64, 0, (hipStream_t)stream>>>((const bf16_t*)A, (const bf16_t*)B, D);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_conv_fwd(const void* x, const void* w, const float* bias, const void* gate, void* y, int B, int H, int W,
                     int cin_s, int cout, int cout_s, int ks, int relu, int out_f32, void* stream) {
  const int Kp = pty_net_kpad(ks, cin_s);
  // fp32 outputs (the network head) may use any channel stride >= cout
  int rc = check_conv_shape(B, H, W, cin_s, cout, out_f32 ? (cout + 7) / 8 * 8 : cout_s, ks, Kp);
  if (rc) return rc;
  if (cout > cout_s) return set_err(PTY_NET_EINVAL, "conv: cout %d > stride %d", cout, cout_s);
  if (gate && out_f32) return set_err(PTY_NET_EINVAL, "conv: a gate needs a bf16 output");
  if (!x || !w || !y) return set_err(PTY_NET_EINVAL, "conv: null pointer");
  ConvFwd a;
  a.x = (const bf16_t*)x;
  a.w = (const bf16_t*)w;
  a.bias = bias;
  a.gate = (const bf16_t*)gate;
  a.y = y;
  a.B = B, a.H = H, a.W = W, a.cin_s = cin_s, a.cout = cout, a.cout_s = cout_s, a.ks = ks;
  a.K = ks * ks * cin_s, a.Kp = Kp, a.relu = relu, a.out_f32 = out_f32;
  const int M = B * H * W;
  dim3 grid((M + CV_BM - 1) / CV_BM, (cout + CV_BN - 1) / CV_BN);
  conv_fwd_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(a);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_conv_wgrad(const void* x, const void* dz, float* part, int splits, float* dw, float* db, int B, int H, int W,
                       int cin_s, int cout, int cout_s, int ks, void* stream) {
  const int Kp = pty_net_kpad(ks, cin_s);
  int rc = check_conv_shape(B, H, W, cin_s, cout, cout_s, ks, Kp);
  if (rc) return rc;
  if (!x || !dz || !part || !dw || splits < 1) return set_err(PTY_NET_EINVAL, "wgrad: bad arguments");
  const int M = B * H * W;
  ConvWgrad a;
  a.x = (const bf16_t*)x;
  a.dz = (const bf16_t*)dz;
  a.part = part;
  a.B = B, a.H = H, a.W = W, a.cin_s = cin_s, a.cout = cout, a.cout_s = cout_s, a.ks = ks;
  a.K = ks * ks * cin_s, a.Kp = Kp;
  int pps = (M + splits - 1) / splits;
  pps = (pps + WG_PT - 1) / WG_PT * WG_PT;
  const int used = (M + pps - 1) / pps;
  a.pix_per_split = pps;
  dim3 grid(Kp / WG_BN + (Kp % WG_BN ? 1 : 0), (cout + WG_BM - 1) / WG_BM, used);
  conv_wgrad_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(a);
  HIP_TRY(hipGetLastError());
  const int n = cout * Kp;
  wgrad_reduce_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(part, used, cout, Kp, a.K, dw, db);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_weight_flip(const void* w, int cin, int cin_s, int cout, int cout_s, int ks, void* wt, void* stream) {
  if (!w || !wt || cin < 1 || cout < 1 || cin > cin_s || cout > cout_s) return set_err(PTY_NET_EINVAL, "weight_flip: bad arguments");
  const int Kp = pty_net_kpad(ks, cin_s), Kpt = pty_net_kpad(ks, cout_s);
  const int n = cin * Kpt;
  weight_flip_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>((const bf16_t*)w, Kp, cin, cin_s, cout, cout_s,
                                                                        ks, (bf16_t*)wt, Kpt);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_pool_fwd(const void* x, void* y, int B, int H, int W, int c_s, void* stream) {
  if (!x || !y || H % 2 || W % 2 || c_s % 8) return set_err(PTY_NET_EINVAL, "pool_fwd: bad arguments");
  pool_fwd_kernel<<<grid_for((size_t)B * H * W * c_s / 4), 256, 0, (hipStream_t)stream>>>((const bf16_t*)x, (bf16_t*)y, B, H, W, c_s);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_pool_bwd(const void* dy, void* dx, int B, int H, int W, int c_s, void* stream) {
  if (!dy || !dx || H % 2 || W % 2 || c_s % 8) return set_err(PTY_NET_EINVAL, "pool_bwd: bad arguments");
  pool_bwd_kernel<<<grid_for((size_t)B * H * W * c_s), 256, 0, (hipStream_t)stream>>>((const bf16_t*)dy, (bf16_t*)dx, B, H, W, c_s);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_up_concat(const void* xa, int ca_s, const void* xb, int cb_s, void* y, int B, int H, int W, void* stream) {
  if (!xa || !xb || !y || H % 2 || W % 2 || ca_s % 8 || cb_s % 8) return set_err(PTY_NET_EINVAL, "up_concat: bad arguments");
  if ((size_t)B * H * W * ((ca_s + cb_s) / 8) >= ((size_t)1 << 31)) return set_err(PTY_NET_EINVAL, "up_concat: too large");
  up_concat_kernel<<<grid_for((size_t)B * H * W * ((ca_s + cb_s) / 8)), 256, 0, (hipStream_t)stream>>>(
      (const bf16_t*)xa, ca_s, (const bf16_t*)xb, cb_s, (bf16_t*)y, B, H, W);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_up_concat_bwd(const void* dy, int ca_s, int cb_s, void* dxa, const void* dxb_in, void* dxb, int B, int H, int W,
                          void* stream) {
  if (!dy || !dxa || !dxb || H % 2 || W % 2 || ca_s % 8 || cb_s % 8) return set_err(PTY_NET_EINVAL, "up_concat_bwd: bad arguments");
  const size_t n = (size_t)B * (H / 2) * (W / 2) * ca_s + (size_t)B * H * W * cb_s;
  up_concat_bwd_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const bf16_t*)dy, ca_s, cb_s, (bf16_t*)dxa,
                                                                     (const bf16_t*)dxb_in, (bf16_t*)dxb, B, H, W);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_add_gate(const void* g1, const void* g2, const void* act, void* y, size_t n, void* stream) {
  if (!g1 || !y) return set_err(PTY_NET_EINVAL, "add_gate: null pointer");
  add_gate_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((const bf16_t*)g1, (const bf16_t*)g2, (const bf16_t*)act,
                                                                (bf16_t*)y, n);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_mse(const float* out, const float* tgt, int B, int Ho, int Wo, int oy, int ox, int hc, int wc, int c_s,
                void* dout, double* partial, int nblocks, void* stream) {
  if (!out || !tgt || !dout || !partial || nblocks < 1 || c_s % 8 || B < 1 || hc < 1 || wc < 1 || oy < 0 || ox < 0 ||
      oy + hc > Ho || ox + wc > Wo)
    return set_err(PTY_NET_EINVAL, "mse: bad arguments");
  mse_kernel<<<nblocks, 256, 0, (hipStream_t)stream>>>(out, tgt, B, Ho, Wo, oy, ox, hc, wc, c_s, (bf16_t*)dout, partial);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_adam(float* w, void* w16, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                 float eps, int step, void* stream) {
  if (!w || !g || !m || !v || step < 1) return set_err(PTY_NET_EINVAL, "adam: bad arguments");
  const float c1 = (float)(1.0 - __builtin_pow((double)b1, step));
  const float c2 = (float)(1.0 - __builtin_pow((double)b2, step));
  adam_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(w, (bf16_t*)w16, g, m, v, n, lr, b1, b2, eps, c1, c2);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

int pty_net_cast_bf16(const float* x, void* y, size_t n, void* stream) {
  if (!x || !y) return set_err(PTY_NET_EINVAL, "cast: null pointer");
  cast_bf16_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>(x, (bf16_t*)y, n);
  HIP_TRY(hipGetLastError());
  return PTY_NET_OK;
}

}  // extern "C"
