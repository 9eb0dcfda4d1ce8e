/*
 * sunlightlb_nn.h -- C ABI of the warm-start network kernels (gfx950 HIP,
 * sunlightlb_amd/csrc/slb_nn.hip).  No counterpart in the reference: the
 * reference has no learned component.  The host network (slb_warm.c) and
 * the kernel tests call these.
 *
 * Tensors are device pointers, channels-last BF16 (uint16_t), X[n][c] with
 * n = esk*(esj*i+j)+k, channel counts multiples of 8.  3x3x3 convolutions
 * with periodic padding; tap t = 9*(di+1) + 3*(dj+1) + (dk+1) reads node
 * (i+di, j+dj, k+dk).  Weights W[co][Kp] with k = t*cin + ci, Kp =
 * slbn_kpad(cin) = 27*cin rounded up to 32, zero-filled (cout rows).  Every
 * call is asynchronous on
 * the stream set by slbn_set_stream (default: the null stream).
 */
#ifndef SUNLIGHTLB_NN_H_INC
#define SUNLIGHTLB_NN_H_INC
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
int  slbn_kpad(int cin);
void slbn_set_stream(void *stream);
/* Y = act(bias + conv3x3x3(X, W)), periodic padding; transposed != 0 uses
 * the mirrored taps (gradient of the input with W = slbn_wtrans(W)).  Y
 * (BF16) and/or Yf (FP32) may be NULL. */
void slbn_conv(const uint16_t *X, int cin, const uint16_t *W, const float *bias, uint16_t *Y,
               float *Yf, int cout, int relu, int transposed, int D, int H, int Wd);
/* dW[co][t*cin+ci] += sum_n dY[n][co] X[n+off(t)][ci]; channels 16..64 step
 * 16.  When slbn_wgrad_lds16_ok() (16 -> 16 channels, x extent a multiple of
 * 8 up to 128, even y extent) bricks of two x-rows are staged in LDS and all
 * 27 taps computed there; else when slbn_wgrad_blocked_ok() (x extent a
 * multiple of 8, channels 16, 32 or 64) it is a GEMM over a node-blocked copy
 * of X and dY; both with a fixed summation order.  Otherwise a per-tap
 * LDS-staged kernel with atomics. */
int  slbn_wgrad_lds16_ok(int cin, int cout, int D, int H, int Wd);
int  slbn_wgrad_blocked_ok(int cin, int cout, int D, int H, int Wd);
void slbn_conv_wgrad(const uint16_t *X, int cin, const uint16_t *dY, int cout, float *dW, int D,
                     int H, int Wd);
/* the same for any channel counts that are multiples of 8 (scalar FP32) */
void slbn_conv_wgrad_small(const uint16_t *X, int cin, const uint16_t *dY, int cout, float *dW,
                           int D, int H, int Wd);
void slbn_wtrans(const uint16_t *W, int cout, int cin, uint16_t *WT);
void slbn_relu_bwd(const uint16_t *Y, const uint16_t *dY, uint16_t *dP, long n);
void slbn_bias_grad(const uint16_t *dP, int C, int N, float *db);
/* dP = dY * (Y > 0) (N nodes x C channels) and db[c] += sum_n dP[n][c] in one
 * pass: slbn_relu_bwd followed by slbn_bias_grad */
void slbn_relu_bwd_bias(const uint16_t *Y, const uint16_t *dY, uint16_t *dP, int C, int N,
                        float *db);
void slbn_pool(const uint16_t *X, uint16_t *Y, int C, int D, int H, int Wd);
void slbn_pool_bwd(const uint16_t *dY, uint16_t *dX, int C, int D, int H, int Wd);
void slbn_up_add(const uint16_t *Xc, const uint16_t *S, uint16_t *Y, int C, int D, int H, int Wd);
void slbn_up_bwd(const uint16_t *dY, uint16_t *dXc, int C, int D, int H, int Wd);
void slbn_add_into(uint16_t *A, const uint16_t *B, long n);
/* loss[0] += sum_n sum_{c<Cv} wc[c] (Yf[n][c] - T[n][c])^2 and dY[n][c] =
 * 2 wc[c] (Yf - T) (0 for c >= Cv); wc is a device array of Cv floats */
void slbn_mse(const float *Yf, int C, const uint16_t *T, int Ct, int Cv, int N, const float *wc,
              uint16_t *dY, float *loss);
void slbn_adam(float *w, uint16_t *wb, const float *g, float *m, float *v, long n, float lr,
               float b1, float b2, float eps, int step);
void slbn_f2bf(const float *a, uint16_t *b, long n);
void slbn_bf2f(const uint16_t *a, float *b, long n);
/* expand a packed sample (Cs channels) into a C-channel tensor, zeros beyond Cs */
void slbn_expand(const uint16_t *S, int Cs, uint16_t *X, int C, long N);
/* device memory (exits with a message on failure) and stream sync */
void *slbn_malloc(size_t bytes);
void slbn_free(void *p);
void slbn_zero(void *p, size_t bytes);
void slbn_h2d(void *dst, const void *src, size_t bytes);
void slbn_d2h(void *dst, const void *src, size_t bytes);
void slbn_sync(void);
#ifdef __cplusplus
}
#endif
#endif
