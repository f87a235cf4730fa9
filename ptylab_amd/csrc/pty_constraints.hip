// Per-iteration constraints of applyConstraints (BaseEngine.py:1440-1612)
// that act on the device state between position loops: the probe power
// normalisation (FP64 reduction), the probe windows, and the centre-of-mass
// stabilisation (FP64 moments, then a roll of probe and object).  Reductions
// run in one 1024-thread workgroup with a fixed summation tree, so results
// are deterministic.
#include <hip/hip_runtime.h>

#include <mutex>

#include "pty_hip.h"

// pty_last_error() message setter of libptyhip (pty_hip.hip), not exported
extern "C" int pty_internal_set_error(int code, const char* msg) __attribute__((visibility("hidden")));

namespace {

constexpr int RED_NT = 1024;

__device__ __forceinline__ double block_reduce(double v, double* red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int s = RED_NT / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

__global__ __launch_bounds__(RED_NT) void power_kernel(const float2* __restrict__ x, size_t n, double* __restrict__ out) {
  __shared__ double red[RED_NT];
  double acc = 0.0;
  for (size_t i = threadIdx.x; i < n; i += RED_NT) {
    const float2 v = x[i];
    acc += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  const double s = block_reduce(acc, red);
  if (threadIdx.x == 0) out[0] = s;
}

__global__ void scale_kernel(float2* __restrict__ x, size_t n, const double* __restrict__ power, float target) {
  const float s = (float)((double)target / sqrt(power[0]));
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float2 v = x[i];
    x[i] = make_float2(v.x * s, v.y * s);
  }
}

// x = (1 - a) x + (a x) w, with w the (N, N) window; a = 1 multiplies by w
__global__ void window_kernel(float2* __restrict__ x, size_t n, size_t nn, const float* __restrict__ w, float a) {
  const float oma = 1.0f - a;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const float2 v = x[i];
    const float wv = w[i % nn];
    if (a == 1.0f) {
      x[i] = make_float2(v.x * wv, v.y * wv);
    } else {
      const float2 av = make_float2(a * v.x, a * v.y);
      x[i] = make_float2(oma * v.x + av.x * wv, oma * v.y + av.y * wv);
    }
  }
}

__global__ __launch_bounds__(RED_NT) void moments_kernel(const float2* __restrict__ x, size_t nn,
                                                         const float* __restrict__ X, const float* __restrict__ Y,
                                                         double* __restrict__ out) {
  __shared__ double red[RED_NT];
  double sx = 0.0, sy = 0.0, s0 = 0.0;
  for (size_t i = threadIdx.x; i < nn; i += RED_NT) {
    const float2 v = x[i];
    const double a = sqrt((double)v.x * (double)v.x + (double)v.y * (double)v.y);
    sx += (double)X[i] * a;
    sy += (double)Y[i] * a;
    s0 += a;
  }
  const double rx = block_reduce(sx, red);
  const double ry = block_reduce(sy, red);
  const double r0 = block_reduce(s0, red);
  if (threadIdx.x == 0) {
    out[0] = rx;
    out[1] = ry;
    out[2] = r0;
  }
}

// dst[f][(y + dy) mod Ny][(x + dx) mod Nx] = src[f][y][x]  (numpy.roll)
__global__ void roll_kernel(const float2* __restrict__ src, float2* __restrict__ dst, size_t nfields, int Ny, int Nx,
                            int dy, int dx) {
  const size_t plane = (size_t)Ny * Nx, n = nfields * plane;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t f = i / plane, r = i - f * plane;
    const int y = (int)(r / Nx), xx = (int)(r - (size_t)y * Nx);
    int ty = (y + dy) % Ny, tx = (xx + dx) % Nx;
    if (ty < 0) ty += Ny;
    if (tx < 0) tx += Nx;
    dst[f * plane + (size_t)ty * Nx + tx] = src[i];
  }
}

__global__ void copy_kernel(const float2* __restrict__ src, float2* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

int grid_for(size_t n) {
  size_t g = (n + 255) / 256;
  if (g > 8192) g = 8192;
  return (int)(g ? g : 1);
}

// Scratch of pty_roll_fields_noalloc: a file-scope device array, one copy
// per GPU, held until the process ends.  Host code never uses its name as an
// address: the device address comes from hipGetSymbolAddress, once per device.
constexpr size_t ROLL_SCRATCH_MAX = PTY_ROLL_SCRATCH_MAX_VALUES;
static __device__ __align__(256) float2 scratch[ROLL_SCRATCH_MAX];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float2* g_scratch_ptr[MAX_DEVICES] = {};

// The calling thread's current device's copy of `scratch`, looked up on the
// first call on that device.
int device_scratch(float2** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= MAX_DEVICES)
    return pty_internal_set_error(PTY_EDEVICE, "pty_roll_fields_noalloc: no current HIP device");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return pty_internal_set_error(PTY_EDEVICE, "pty_roll_fields_noalloc: hipGetSymbolAddress(scratch) failed");
    g_scratch_ptr[dev] = (float2*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return PTY_OK;
}

}  // namespace

extern "C" {

int pty_field_power(const void* x, size_t n, double* out, void* stream) {
  if (!x || !out || n == 0) return pty_internal_set_error(PTY_EINVAL, "pty_field_power: bad arguments");
  power_kernel<<<1, RED_NT, 0, (hipStream_t)stream>>>((const float2*)x, n, out);
  return hipGetLastError() == hipSuccess ? PTY_OK : pty_internal_set_error(PTY_EDEVICE, "constraint kernel launch failed");
}

int pty_scale_to_power(void* x, size_t n, const double* power, float target, void* stream) {
  if (!x || !power || n == 0) return pty_internal_set_error(PTY_EINVAL, "pty_scale_to_power: bad arguments");
  scale_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((float2*)x, n, power, target);
  return hipGetLastError() == hipSuccess ? PTY_OK : pty_internal_set_error(PTY_EDEVICE, "constraint kernel launch failed");
}

int pty_window_fields(void* x, int32_t nfields, int32_t N, const float* window, float aleph, void* stream) {
  if (!x || !window || nfields < 1 || N < 1) return pty_internal_set_error(PTY_EINVAL, "pty_window_fields: bad arguments");
  const size_t nn = (size_t)N * N, n = (size_t)nfields * nn;
  window_kernel<<<grid_for(n), 256, 0, (hipStream_t)stream>>>((float2*)x, n, nn, window, aleph);
  return hipGetLastError() == hipSuccess ? PTY_OK : pty_internal_set_error(PTY_EDEVICE, "constraint kernel launch failed");
}

int pty_field_moments(const void* x, int32_t N, const float* X, const float* Y, double* out3, void* stream) {
  if (!x || !X || !Y || !out3 || N < 1) return pty_internal_set_error(PTY_EINVAL, "pty_field_moments: bad arguments");
  moments_kernel<<<1, RED_NT, 0, (hipStream_t)stream>>>((const float2*)x, (size_t)N * N, X, Y, out3);
  return hipGetLastError() == hipSuccess ? PTY_OK : pty_internal_set_error(PTY_EDEVICE, "constraint kernel launch failed");
}

int pty_roll_fields(void* x, void* scratch, int32_t nfields, int32_t Ny, int32_t Nx, int32_t dy, int32_t dx,
                    void* stream) {
  if (!x || !scratch || nfields < 1 || Ny < 1 || Nx < 1)
    return pty_internal_set_error(PTY_EINVAL, "pty_roll_fields: bad arguments");
  const size_t n = (size_t)nfields * Ny * Nx;
  hipStream_t st = (hipStream_t)stream;
  roll_kernel<<<grid_for(n), 256, 0, st>>>((const float2*)x, (float2*)scratch, (size_t)nfields, Ny, Nx, dy, dx);
  if (hipGetLastError() != hipSuccess) return pty_internal_set_error(PTY_EDEVICE, "pty_roll_fields: launch failed");
  if (hipMemcpyAsync(x, scratch, n * sizeof(float2), hipMemcpyDeviceToDevice, st) != hipSuccess)
    return pty_internal_set_error(PTY_EDEVICE, "pty_roll_fields: copy back failed");
  return PTY_OK;
}

int pty_roll_fields_noalloc(void* x, int32_t nfields, int32_t Ny, int32_t Nx, int32_t dy, int32_t dx, void* stream) {
  if (!x || nfields < 1 || Ny < 1 || Nx < 1)
    return pty_internal_set_error(PTY_EINVAL, "pty_roll_fields_noalloc: bad arguments");
  const size_t n = (size_t)nfields * Ny * Nx;
  if (n > ROLL_SCRATCH_MAX)
    return pty_internal_set_error(PTY_EINVAL, "pty_roll_fields_noalloc: more than PTY_ROLL_SCRATCH_MAX_VALUES "
                                              "complex values (use pty_roll_fields with a workspace)");
  float2* buf = nullptr;
  const int rc = device_scratch(&buf);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  // the roll writes every element of buf[0, n) that the copy reads: nothing to clear
  roll_kernel<<<grid_for(n), 256, 0, st>>>((const float2*)x, buf, (size_t)nfields, Ny, Nx, dy, dx);
  if (hipGetLastError() != hipSuccess) return pty_internal_set_error(PTY_EDEVICE, "pty_roll_fields_noalloc: launch failed");
  copy_kernel<<<grid_for(n), 256, 0, st>>>(buf, (float2*)x, n);
  if (hipGetLastError() != hipSuccess) return pty_internal_set_error(PTY_EDEVICE, "pty_roll_fields_noalloc: launch failed");
  return PTY_OK;
}

}  // extern "C"
