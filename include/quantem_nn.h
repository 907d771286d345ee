/*
 * quantem_nn.h -- C ABI of the learned aberration initializer's network
 * runtime in libquantem_ssb_hip.so (gfx950).  These entry points have no
 * counterpart in the reference: the network only seeds the batched
 * aberration search, whose exact SSB objective scores its output.
 *
 * All arrays are device pointers; `stream` is a hipStream_t.  Status 0 on
 * success, -1 for bad arguments, -4 for a launch failure.  bf16 values are
 * passed as raw 16-bit patterns (uint16_t).
 */
#ifndef QUANTEM_NN_H
#define QUANTEM_NN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* C(m,n) (=, or += when accumulate) sum_k A(m,k) B(k,n) [+ bias(n)] [relu]:
 * A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk + n*sbn] (bf16), C fp32 at
 * C[m*scm + n*scn]; fp32 accumulation on the matrix cores. */
int qnn_gemm_bf16(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk, long long sbn,
                  float* C, long long scm, long long scn, const float* bias, int M, int N, int K, int relu,
                  int accumulate, void* stream);

/* The same product, row-major C[M][N] (= only), with K split into up to
 * `splits` slices of whole 16-steps computed in parallel into workspace
 * (splits * M * N floats) and summed in slice order. */
int qnn_gemm_bf16_splitk(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk,
                         long long sbn, float* C, int M, int N, int K, int splits, float* workspace, void* stream);

/* NHWC fp32 [B][H][W][Cin] -> bf16 columns [(b,oy,ox)][(ky,kx,c)] of a 3x3,
 * stride 2, pad 1 convolution (H, W even). */
int qnn_im2col(const float* X, int B, int H, int W, int Cin, uint16_t* cols, void* stream);

/* Adjoint of qnn_im2col on fp32 column gradients (a fixed-order gather). */
int qnn_col2im(const float* dcols, int B, int H, int W, int Cin, float* dX, void* stream);

/* fp32 -> bf16, round to nearest even. */
int qnn_cast_bf16(const float* x, long long n, uint16_t* y, void* stream);

/* dZ = dY * (Y > 0) (Y NULL: identity), with its bf16 copy dZh. */
int qnn_relu_bwd(const float* Y, const float* dY, long long n, float* dZ, uint16_t* dZh, void* stream);

/* out[n] = sum_m dZ[m][n] (compensated, fixed order: partial sums over
 * chunks of rows, then over the chunks).  `work` holds at least
 * qnn_colsum_workspace(M, N) floats of device memory. */
long long qnn_colsum_workspace(int M, int N);
int qnn_colsum(const float* dZ, int M, int N, float* out, float* work, long long work_floats, void* stream);

/* loss = mean (pred - target)^2 over n values; grad = 2 (pred - target) / n. */
int qnn_mse(const float* pred, const float* target, int n, float* grad, float* loss, void* stream);

/* One Adam step (bias-corrected, step >= 1) on fp32 parameters; ph, if not
 * NULL, receives the bf16 copy. */
int qnn_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
             int step, uint16_t* ph, void* stream);

/* Forward-model training features: S samples theta[s] = (C10 A, C30 A,
 * rotation rad), noise sigma snr[s], detector reciprocal sampling recip
 * (1/A per pixel), semiangle and angular sampling (rad), wavelength (A),
 * q_host[16] = 8 (qx, qy) scan frequencies (1/A), generator seed;
 * out [S][32][32][16] NHWC. */
int qnn_features_synth(const float* theta, const float* snr, int S, float recip, float semi_rad, float ang_rad,
                       float wavelength, const float* q_host, unsigned int seed, float* out, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* QUANTEM_NN_H */
