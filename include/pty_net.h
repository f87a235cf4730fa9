/*
 * pty_net.h -- C ABI of the learned initialiser's gfx950 kernels
 * (libptynet.so): the layers of a small U-Net, their backward passes, an MSE
 * loss and Adam, on BF16 matrix cores with FP32 accumulation.
 *
 * The reference has no learned initialiser: its initial object is the
 * 'ones'/'circ' guess of initialProbeOrObject
 * (utils/initializationFunctions.py:10-62), set through
 * Reconstruction.initializeObject (Reconstruction.py:356-374).  These entry
 * points build the network whose stitched output replaces that guess; none
 * replaces a reference function one-to-one.
 *
 * Tensors are device pointers, NHWC, bf16 (raw 16-bit) unless stated, with a
 * channel stride (c_s) that is a multiple of 8 and zeros in the padding
 * channels.  A convolution's weights are (cout, Kp) bf16 rows in
 * (ky, kx, ci) order, ci running over the input channel stride, padded with
 * zeros to Kp = pty_net_kpad(ks, cin_s).  Every function returns 0 or a
 * negative PTY_NET_E* code (message in pty_net_last_error()); all are
 * stream-ordered and allocate nothing.
 */
#ifndef PTY_NET_H
#define PTY_NET_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PTY_NET_OK 0
#define PTY_NET_EINVAL (-22)
#define PTY_NET_EDEVICE (-5)

const char* pty_net_last_error(void);

/* Weight row length of a ks x ks convolution over cin_s input channels:
 * ks*ks*cin_s rounded up (with one spare column) to a multiple of 32. */
int pty_net_kpad(int ks, int cin_s);

/* One v_mfma_f32_32x32x16_bf16 on A (32x16) and B (16x32) bf16 row-major,
 * D (32x32) fp32: checks the operand layout the kernels assume. */
int pty_net_mfma_probe(const void* A, const void* B, float* D, void* stream);

/* y = act(conv_ks(x, w) + bias) [* (gate > 0)], 'same' zero padding.
 * x (B,H,W,cin_s); y (B,H,W,cout_s) bf16, or fp32 when out_f32 != 0;
 * act = ReLU when relu != 0; gate (B,H,W,cout_s) or NULL. */
int pty_net_conv_fwd(const void* x, const void* w, const float* bias, const void* gate, void* y, int B, int H, int W,
                     int cin_s, int cout, int cout_s, int ks, int relu, int out_f32, void* stream);

/* Weight and bias gradients of a convolution: dw (cout, Kp) fp32 and
 * db (cout) fp32 (NULL: skipped) from the layer input x (B,H,W,cin_s) and
 * the pre-activation gradient dz (B,H,W,cout_s); the pixel sum is split in
 * `splits` ranges with fp32 partials in `part` (splits*cout*Kp floats) and
 * reduced in a fixed order. */
int pty_net_conv_wgrad(const void* x, const void* dz, float* part, int splits, float* dw, float* db, int B, int H, int W,
                       int cin_s, int cout, int cout_s, int ks, void* stream);

/* Weights of the input-gradient convolution: wt (cin, pty_net_kpad(ks,
 * cout_s)) with wt[ci, (ky,kx,co)] = w[co, (ks-1-ky, ks-1-kx, ci)]. */
int pty_net_weight_flip(const void* w, int cin, int cin_s, int cout, int cout_s, int ks, void* wt, void* stream);

/* 2x2 average pooling (B,H,W,c_s) -> (B,H/2,W/2,c_s), and its backward. */
int pty_net_pool_fwd(const void* x, void* y, int B, int H, int W, int c_s, void* stream);
int pty_net_pool_bwd(const void* dy, void* dx, int B, int H, int W, int c_s, void* stream);

/* y (B,H,W,ca_s+cb_s) = [nearest 2x upsampling of xa (B,H/2,W/2,ca_s) | xb
 * (B,H,W,cb_s)], and its backward (dxb = skip part [+ dxb_in]). */
int pty_net_up_concat(const void* xa, int ca_s, const void* xb, int cb_s, void* y, int B, int H, int W, void* stream);
int pty_net_up_concat_bwd(const void* dy, int ca_s, int cb_s, void* dxa, const void* dxb_in, void* dxb, int B, int H,
                          int W, void* stream);

/* y = (g1 [+ g2]) * (act > 0 if act != NULL): sums two gradient
 * contributions at a ReLU output and applies the ReLU derivative. */
int pty_net_add_gate(const void* g1, const void* g2, const void* act, void* y, size_t n, void* stream);

/* Mean squared error of the crop [oy, oy+hc) x [ox, ox+wc) of the fp32
 * outputs out (B,Ho,Wo,2) (real, imaginary) against tgt (B,hc,wc,2):
 * dout (B,Ho,Wo,c_s) bf16 = 2 (out - tgt) / n inside the crop
 * (n = B*hc*wc*2; entries outside are not written); partial[b] is the FP64
 * sum of squares of block b (nblocks blocks). */
int pty_net_mse(const float* out, const float* tgt, int B, int Ho, int Wo, int oy, int ox, int hc, int wc, int c_s,
                void* dout, double* partial, int nblocks, void* stream);

/* One Adam step (step >= 1) on n fp32 weights; refreshes the bf16 copy w16
 * (NULL: none). */
int pty_net_adam(float* w, void* w16, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2,
                 float eps, int step, void* stream);

/* y (bf16) = x (fp32), round to nearest even. */
int pty_net_cast_bf16(const float* x, void* y, size_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* PTY_NET_H */
