// Learned aberration initializer: the device side of a small convolutional
// network trained and run on the GPU.  No reference counterpart exists; the
// network predicts (C10, C30, rotation) from trotter feature maps and only
// seeds the batched search, whose exact SSB objective scores the result.
//
// Every layer is a GEMM on the bf16 matrix cores (inputs rounded to bf16,
// products accumulated in fp32), fed by im2col for the 3x3 stride-2
// convolutions; activations are NHWC fp32.  Backward passes reuse the same
// GEMM with transposed strides; the only non-GEMM reductions (bias
// gradients, col2im) are gathers in a fixed order, so a training run is
// bit-repeatable.
//
// Also here: the synthetic training features (trotter values of the SSB
// forward model G_b(Q) = Gamma_b(Q; theta) Phi(Q) + noise at fixed Q) and
// the same features read from a resident Fourier stack.
#include "ssb_internal.h"
#include "ssb_overlap.h"

namespace qssb {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f32_to_bf16(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);   // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                              // round to nearest even
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// ------------------------------------------------------------------------
// C(m, n) (=|+=) sum_k A(m, k) B(k, n) [+ bias(n)] [relu], A and B bf16 with
// arbitrary strides, C fp32.  64 x 64 block tile, 4 waves of 32 x 32, each
// a 2 x 2 grid of 16x16x16 bf16 MFMAs; K in steps of 16 through LDS.
// MFMA operand layout (16x16x16, wave64): lane l supplies A rows l%16 and
// B columns l%16 at k = 4*(l/16) .. +3; accumulator r of lane l holds
// C(4*(l/16) + r, l%16).
// ------------------------------------------------------------------------
template <bool RELU>
__global__ void __launch_bounds__(256)
k_gemm_bf16(const uint16_t* __restrict__ A, long long sam, long long sak, const uint16_t* __restrict__ B,
            long long sbk, long long sbn, float* __restrict__ C, long long scm, long long scn,
            const float* __restrict__ bias, int M, int N, int K, int accumulate, int kslice, long long pstride) {
  __shared__ uint16_t As[64][20];
  __shared__ uint16_t Bs[16][68];
  const int tid = threadIdx.x;
  const int lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  float4v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = float4v{0.f, 0.f, 0.f, 0.f};
  // split K: block z covers [kb, ke) and writes partial plane z
  const int kb = blockIdx.z * kslice;
  const int ke = min(K, kb + kslice);
  C += (long long)blockIdx.z * pstride;
  for (int k0 = kb; k0 < ke; k0 += 16) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int idx = tid + r * 256;
      const int mm = idx >> 4, kk = idx & 15;
      const int gm = m0 + mm, gk = k0 + kk;
      As[mm][kk] = (gm < M && gk < ke) ? A[(long long)gm * sam + (long long)gk * sak] : (uint16_t)0;
      const int kr = idx >> 6, nn = idx & 63;
      const int gk2 = k0 + kr, gn = n0 + nn;
      Bs[kr][nn] = (gk2 < ke && gn < N) ? B[(long long)gk2 * sbk + (long long)gn * sbn] : (uint16_t)0;
    }
    __syncthreads();
    const int kq = (lane >> 4) * 4;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ar = wm + i * 16 + (lane & 15);
      const short4v a = short4v{(short)As[ar][kq], (short)As[ar][kq + 1], (short)As[ar][kq + 2],
                                (short)As[ar][kq + 3]};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int bc = wn + j * 16 + (lane & 15);
        const short4v b = short4v{(short)Bs[kq][bc], (short)Bs[kq + 1][bc], (short)Bs[kq + 2][bc],
                                  (short)Bs[kq + 3][bc]};
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        const int gn = n0 + wn + j * 16 + (lane & 15);
        if (gm < M && gn < N) {
          float v = acc[i][j][r];
          if (bias) v += bias[gn];
          if (RELU) v = fmaxf(v, 0.0f);
          float* c = C + (long long)gm * scm + (long long)gn * scn;
          *c = accumulate ? *c + v : v;
        }
      }
}

// C[m][n] = sum_z P[z][m][n] in order (split-K partial planes, contiguous M x N)
__global__ void __launch_bounds__(256)
k_sum_partials(const float* __restrict__ P, int splits, long long n, float* __restrict__ C) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += P[(long long)z * n + i];
    C[i] = s;
  }
}

// NHWC fp32 -> im2col bf16 for a 3x3, stride 2, pad 1 convolution:
// cols[(b, oy, ox)][(ky, kx, c)] = X[b][2 oy + ky - 1][2 ox + kx - 1][c] (0 outside).
__global__ void __launch_bounds__(256)
k_im2col_bf16(const float* __restrict__ X, int B, int H, int W, int Cin, int Ho, int Wo, uint16_t* __restrict__ cols) {
  const long long K = 9LL * Cin;
  const long long total = (long long)B * Ho * Wo * K;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const long long m = idx / K;
    const int k = (int)(idx - m * K);
    const int c = k % Cin, kxy = k / Cin, kx = kxy % 3, ky = kxy / 3;
    const int ox = (int)(m % Wo), oy = (int)((m / Wo) % Ho), b = (int)(m / ((long long)Wo * Ho));
    const int y = 2 * oy + ky - 1, x = 2 * ox + kx - 1;
    float v = 0.0f;
    if (y >= 0 && y < H && x >= 0 && x < W) v = X[(((long long)b * H + y) * W + x) * Cin + c];
    cols[idx] = f32_to_bf16(v);
  }
}

// dX[b][y][x][c] = sum over the output positions whose 3x3 window covers
// (y, x) of dcols[(b, oy, ox)][(ky, kx, c)], in (ky, kx) order.
__global__ void __launch_bounds__(256)
k_col2im(const float* __restrict__ dcols, int B, int H, int W, int Cin, int Ho, int Wo, float* __restrict__ dX) {
  const long long total = (long long)B * H * W * Cin;
  for (long long idx = (long long)blockIdx.x * 256 + threadIdx.x; idx < total; idx += (long long)gridDim.x * 256) {
    const int c = (int)(idx % Cin);
    const int x = (int)((idx / Cin) % W), y = (int)((idx / ((long long)Cin * W)) % H);
    const int b = (int)(idx / ((long long)Cin * W * H));
    float s = 0.0f;
    for (int ky = 0; ky < 3; ++ky) {
      const int ty = y + 1 - ky;
      if (ty < 0 || (ty & 1) || (ty >> 1) >= Ho) continue;
      for (int kx = 0; kx < 3; ++kx) {
        const int tx = x + 1 - kx;
        if (tx < 0 || (tx & 1) || (tx >> 1) >= Wo) continue;
        const long long m = ((long long)b * Ho + (ty >> 1)) * Wo + (tx >> 1);
        s += dcols[m * 9 * Cin + (ky * 3 + kx) * Cin + c];
      }
    }
    dX[idx] = s;
  }
}

__global__ void __launch_bounds__(256)
k_cast_bf16(const float* __restrict__ x, long long n, uint16_t* __restrict__ y) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    y[i] = f32_to_bf16(x[i]);
}

// dZ = dY * (Y > 0), also rounded to bf16 for the backward GEMMs
__global__ void __launch_bounds__(256)
k_relu_bwd(const float* __restrict__ Y, const float* __restrict__ dY, long long n, float* __restrict__ dZ,
           uint16_t* __restrict__ dZh) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float g = (Y == nullptr || Y[i] > 0.0f) ? dY[i] : 0.0f;
    dZ[i] = g;
    dZh[i] = f32_to_bf16(g);
  }
}

// bias gradient: column sums of dZ[M][N] in a fixed order, in two passes.
// Pass 1: one workgroup per (64-column tile, chunk of kColsumRows rows),
// 4 row lanes per column, each a compensated sum, combined in lane order;
// pass 2: the chunks' partial sums per column, compensated, in chunk order.
constexpr int kColsumRows = 2048;

__global__ void __launch_bounds__(256)
k_colsum_part(const float* __restrict__ dZ, int M, int N, float* __restrict__ part) {
  __shared__ float sh[4][64];
  const int lc = threadIdx.x & 63, lr = threadIdx.x >> 6;
  const int n = blockIdx.x * 64 + lc;
  const int m0 = blockIdx.y * kColsumRows;
  const int m1 = min(M, m0 + kColsumRows);
  float s = 0.0f, c = 0.0f;
  if (n < N) {
    for (int m = m0 + lr; m < m1; m += 4) {
      const float y = dZ[(long long)m * N + n] - c;
      const float t = s + y;
      c = (t - s) - y;
      s = t;
    }
  }
  sh[lr][lc] = s - c;
  __syncthreads();
  if (lr == 0 && n < N) {
    float s2 = 0.0f, c2 = 0.0f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float y = sh[r][lc] - c2;
      const float t = s2 + y;
      c2 = (t - s2) - y;
      s2 = t;
    }
    part[(long long)blockIdx.y * N + n] = s2 - c2;
  }
}

__global__ void __launch_bounds__(256)
k_colsum_finish(const float* __restrict__ part, int chunks, int N, float* __restrict__ out) {
  const int n = blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f, c = 0.0f;
  for (int k = 0; k < chunks; ++k) {
    const float y = part[(long long)k * N + n] - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  out[n] = s - c;
}

// mean squared error over B x D and its gradient (2/(B D)) (p - t)
__global__ void __launch_bounds__(256)
k_mse(const float* __restrict__ pred, const float* __restrict__ target, int n, float* __restrict__ grad,
      float* __restrict__ loss) {
  __shared__ float sm[256];
  float s = 0.0f;
  for (int i = threadIdx.x; i < n; i += 256) {
    const float d = pred[i] - target[i];
    grad[i] = 2.0f * d / (float)n;
    s += d * d;
  }
  sm[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) sm[threadIdx.x] += sm[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = sm[0] / (float)n;
}

// Adam on fp32 master weights; refreshes the bf16 copy
__global__ void __launch_bounds__(256)
k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v, long long n,
       float lr, float b1, float b2, float eps, float bc1, float bc2, uint16_t* __restrict__ ph) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float pi = p[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    p[i] = pi;
    if (ph) ph[i] = f32_to_bf16(pi);
  }
}

// ------------------------------------------------------------------------
// Trotter features.  For each sample and each of NQ fixed scan frequencies
// Q_i, the values G(Q_i, k) over a WxW detector window centred on the BF
// disk become two channels: |G| / max_k |G| and the phase of
// G(k) conj(G(k_c)) / pi, k_c the window centre (the object's phase Phi(Q)
// cancels).  Output NHWC [sample][W][W][2 NQ].
//
// Synthetic samples follow the weak-phase SSB forward model in the engine's
// own conventions: G_b(Q) = Gamma_b(Q; C10, C30, rotation) Phi(Q) plus
// complex Gaussian noise, Gamma from the full chi (C10 and C30 terms) and
// the soft aperture, k rotated by -rotation as in the geometry kernel.
// Random numbers come from a counter-based hash of (seed, sample, index).
// ------------------------------------------------------------------------
__device__ __forceinline__ uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ float uniform01(uint32_t seed, uint32_t a, uint32_t b) {
  const uint32_t h = hash32(seed ^ hash32(a * 0x9e3779b9u ^ hash32(b + 0x632be5abu)));
  return ((float)(h >> 8) + 0.5f) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float2 normal2(uint32_t seed, uint32_t a, uint32_t b) {
  const float u1 = uniform01(seed, a, 2 * b), u2 = uniform01(seed, a, 2 * b + 1);
  const float r = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincosf(6.283185307179586f * u2, &s, &c);
  return make_float2(r * c, r * s);
}

struct FeatureOptics {
  float recip;        // detector reciprocal sampling (1/A per pixel)
  float semi;         // semiangle (rad)
  float ang;          // angular sampling (rad per pixel)
  float lam;          // wavelength (A)
  float qx[8], qy[8]; // the fixed scan frequencies (1/A)
};

__device__ __forceinline__ float soft_aperture(float ux, float uy, const FeatureOptics& f) {
  const float r = sqrtf(ux * ux + uy * uy);
  return fminf(fmaxf((f.semi - f.lam * r) / f.ang + 0.5f, 0.0f), 1.0f);
}

__device__ __forceinline__ float chi_c10_c30(float ux, float uy, float c10, float c30, float lam) {
  const float r2 = ux * ux + uy * uy;
  return 3.14159265358979f * lam * r2 * c10 + 1.5707963267949f * lam * lam * lam * r2 * r2 * c30;
}

template <int NQ, int W>
__global__ void __launch_bounds__(256)
k_features_synth(const float* __restrict__ theta, const float* __restrict__ snr, int S, FeatureOptics f,
                 uint32_t seed, float* __restrict__ out) {
  constexpr int C = 2 * NQ;
  __shared__ float2 g[W * W];
  __shared__ float red[256];
  for (int s = blockIdx.x; s < S; s += gridDim.x) {
    const float c10 = theta[3 * s], c30 = theta[3 * s + 1], rot = theta[3 * s + 2];
    const float cr = cosf(-rot), sr = sinf(-rot);
    // sub-pixel offset of the true disk centre from the window centre
    const float oy = uniform01(seed ^ 0x27d4eb2fu, (uint32_t)s, 1u) - 0.5f;
    const float ox = uniform01(seed ^ 0x27d4eb2fu, (uint32_t)s, 2u) - 0.5f;
    for (int qi = 0; qi < NQ; ++qi) {
      const float qx = f.qx[qi], qy = f.qy[qi];
      const float2 phi = normal2(seed, (uint32_t)s, 1000u + qi);          // object spectrum Phi(Q)
      float local_max = 0.0f;
      for (int p = threadIdx.x; p < W * W; p += 256) {
        const int iy = p / W, ix = p % W;
        const float kxa = ((float)(iy - W / 2) + oy) * f.recip, kya = ((float)(ix - W / 2) + ox) * f.recip;
        const float kx = kxa * cr + kya * sr, ky = -kxa * sr + kya * cr;
        const float ak = soft_aperture(kx, ky, f);
        float2 val = make_float2(0.f, 0.f);
        if (ak > 0.0f) {
          const float am = soft_aperture(qx - kx, qy - ky, f), ap = soft_aperture(qx + kx, qy + ky, f);
          const float chik = chi_c10_c30(kx, ky, c10, c30, f.lam);
          const float dm = chi_c10_c30(qx - kx, qy - ky, c10, c30, f.lam) - chik;
          const float dp = chi_c10_c30(qx + kx, qy + ky, c10, c30, f.lam) - chik;
          float sm, cm, sp, cp;
          sincosf(dm, &sm, &cm);
          sincosf(dp, &sp, &cp);
          const float gr = ak * (am * cm - ap * cp), gi = ak * (-am * sm - ap * sp);
          const float2 nz = normal2(seed ^ 0x5bd1e995u, (uint32_t)s, (uint32_t)(qi * W * W + p));
          val = make_float2(gr * phi.x - gi * phi.y + snr[s] * nz.x, gr * phi.y + gi * phi.x + snr[s] * nz.y);
        }
        g[p] = val;
        local_max = fmaxf(local_max, sqrtf(val.x * val.x + val.y * val.y));
      }
      red[threadIdx.x] = local_max;
      __syncthreads();
      for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
      }
      const float inv = red[0] > 0.0f ? 1.0f / red[0] : 0.0f;
      const float2 gc = g[(W / 2) * W + W / 2];
      for (int p = threadIdx.x; p < W * W; p += 256) {
        const float2 v = g[p];
        const float mag = sqrtf(v.x * v.x + v.y * v.y);
        const float re = v.x * gc.x + v.y * gc.y, im = v.y * gc.x - v.x * gc.y;
        float* o = out + ((long long)s * W * W + p) * C;
        o[2 * qi] = mag * inv;
        o[2 * qi + 1] = (mag > 0.0f) ? atan2f(im, re) * 0.3183098861837907f : 0.0f;
      }
      __syncthreads();
    }
  }
}

// The same features from a resident Hermitian stack Gt[b][c][q] (columns
// 0..N/2; column c > N/2 read as conj(G(-q, N-c))), scattered by the BF
// pixels' detector offsets from the rounded disk centre.
template <int NQ, int W>
__global__ void __launch_bounds__(256)
k_features_from_g(const float2* __restrict__ Gt, int N, const int* __restrict__ bf_rows,
                  const int* __restrict__ bf_cols, const BfInfo* __restrict__ info, const int* __restrict__ qrow,
                  const int* __restrict__ qcol, float* __restrict__ out) {
  constexpr int C = 2 * NQ;
  __shared__ float2 g[W * W];
  __shared__ float red[256];
  const int nbf = info->nbf;
  const int crow = (int)rintf(info->center_r), ccol = (int)rintf(info->center_c);
  const int half = N / 2, gcols = half + 1;
  for (int qi = 0; qi < NQ; ++qi) {
    for (int p = threadIdx.x; p < W * W; p += 256) g[p] = make_float2(0.f, 0.f);
    __syncthreads();
    const int qr = qrow[qi], qc = qcol[qi];
    for (int b = threadIdx.x; b < nbf; b += 256) {
      const int iy = bf_rows[b] - crow + W / 2, ix = bf_cols[b] - ccol + W / 2;
      if (iy < 0 || iy >= W || ix < 0 || ix >= W) continue;
      float2 v;
      if (qc <= half) {
        v = Gt[((long long)b * gcols + qc) * N + qr];
      } else {
        const float2 m = Gt[((long long)b * gcols + (N - qc)) * N + ((N - qr) & (N - 1))];
        v = make_float2(m.x, -m.y);
      }
      g[iy * W + ix] = v;
    }
    __syncthreads();
    float local_max = 0.0f;
    for (int p = threadIdx.x; p < W * W; p += 256) {
      const float2 v = g[p];
      local_max = fmaxf(local_max, sqrtf(v.x * v.x + v.y * v.y));
    }
    red[threadIdx.x] = local_max;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (threadIdx.x < h) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + h]);
      __syncthreads();
    }
    const float inv = red[0] > 0.0f ? 1.0f / red[0] : 0.0f;
    const float2 gc = g[(W / 2) * W + W / 2];
    for (int p = threadIdx.x; p < W * W; p += 256) {
      const float2 v = g[p];
      const float mag = sqrtf(v.x * v.x + v.y * v.y);
      const float re = v.x * gc.x + v.y * gc.y, im = v.y * gc.x - v.x * gc.y;
      out[(long long)p * C + 2 * qi] = mag * inv;
      out[(long long)p * C + 2 * qi + 1] = (mag > 0.0f) ? atan2f(im, re) * 0.3183098861837907f : 0.0f;
    }
    __syncthreads();
  }
}

hipError_t launch_features_from_g(const float2* Gt, int N, const int* bf_rows, const int* bf_cols,
                                  const BfInfo* info, const int* qrow, const int* qcol, float* out, hipStream_t s) {
  hipLaunchKernelGGL((k_features_from_g<8, 32>), dim3(1), dim3(256), 0, s, Gt, N, bf_rows, bf_cols, info, qrow, qcol,
                     out);
  return hipGetLastError();
}

static unsigned grid_of(long long n) {
  long long g = (n + 255) / 256;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (unsigned)g;
}

}  // namespace qssb

// ------------------------------------------------------------------------
// C ABI of the network runtime (device pointers; hipStream_t).
// ------------------------------------------------------------------------
using namespace qssb;

extern "C" {

int qnn_gemm_bf16(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk, long long sbn,
                  float* C, long long scm, long long scn, const float* bias, int M, int N, int K, int relu,
                  int accumulate, void* stream) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1) return -1;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (relu)
    hipLaunchKernelGGL(k_gemm_bf16<true>, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, scm, scn, bias, M, N,
                       K, accumulate, K, 0LL);
  else
    hipLaunchKernelGGL(k_gemm_bf16<false>, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, C, scm, scn, bias, M, N,
                       K, accumulate, K, 0LL);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_gemm_bf16_splitk(const uint16_t* A, long long sam, long long sak, const uint16_t* B, long long sbk,
                         long long sbn, float* C, int M, int N, int K, int splits, float* workspace, void* stream) {
  if (!A || !B || !C || !workspace || M < 1 || N < 1 || K < 1 || splits < 1) return -1;
  const int kslice = ((K + splits - 1) / splits + 15) / 16 * 16;
  const int z = (K + kslice - 1) / kslice;
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64), (unsigned)z);
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  const long long plane = (long long)M * N;
  hipLaunchKernelGGL(k_gemm_bf16<false>, grid, dim3(256), 0, s, A, sam, sak, B, sbk, sbn, workspace, (long long)N,
                     1LL, nullptr, M, N, K, 0, kslice, plane);
  hipLaunchKernelGGL(k_sum_partials, dim3(grid_of(plane)), dim3(256), 0, s, workspace, z, plane, C);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_im2col(const float* X, int B, int H, int W, int Cin, uint16_t* cols, void* stream) {
  if (!X || !cols || B < 1 || H < 2 || W < 2 || Cin < 1) return -1;
  const int Ho = H / 2, Wo = W / 2;
  hipLaunchKernelGGL(k_im2col_bf16, dim3(grid_of((long long)B * Ho * Wo * 9 * Cin)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), X, B, H, W, Cin, Ho, Wo, cols);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_col2im(const float* dcols, int B, int H, int W, int Cin, float* dX, void* stream) {
  if (!dcols || !dX || B < 1 || H < 2 || W < 2 || Cin < 1) return -1;
  hipLaunchKernelGGL(k_col2im, dim3(grid_of((long long)B * H * W * Cin)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), dcols, B, H, W, Cin, H / 2, W / 2, dX);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_cast_bf16(const float* x, long long n, uint16_t* y, void* stream) {
  if (!x || !y || n < 1) return -1;
  hipLaunchKernelGGL(k_cast_bf16, dim3(grid_of(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), x, n, y);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_relu_bwd(const float* Y, const float* dY, long long n, float* dZ, uint16_t* dZh, void* stream) {
  if (!dY || !dZ || !dZh || n < 1) return -1;
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid_of(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), Y, dY, n, dZ,
                     dZh);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

long long qnn_colsum_workspace(int M, int N) {
  if (M < 1 || N < 1) return -1;
  return (long long)((M + kColsumRows - 1) / kColsumRows) * N;
}

int qnn_colsum(const float* dZ, int M, int N, float* out, float* work, long long work_floats, void* stream) {
  if (!dZ || !out || !work || M < 1 || N < 1) return -1;
  const int chunks = (M + kColsumRows - 1) / kColsumRows;
  if (chunks > 65535 || work_floats < (long long)chunks * N) return -1;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_colsum_part, dim3((unsigned)((N + 63) / 64), (unsigned)chunks), dim3(256), 0, s, dZ, M, N,
                     work);
  hipLaunchKernelGGL(k_colsum_finish, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, s, work, chunks, N, out);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_mse(const float* pred, const float* target, int n, float* grad, float* loss, void* stream) {
  if (!pred || !target || !grad || !loss || n < 1) return -1;
  hipLaunchKernelGGL(k_mse, dim3(1), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), pred, target, n, grad, loss);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

int qnn_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
             int step, uint16_t* ph, void* stream) {
  if (!p || !g || !m || !v || n < 1 || step < 1) return -1;
  const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid_of(n)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), p, g, m, v, n, lr,
                     b1, b2, eps, bc1, bc2, ph);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

// theta: [S][3] (C10 A, C30 A, rotation rad); snr: [S] noise sigma (relative to |Phi| ~ 1);
// q_host: [2*8] (qx, qy pairs of the 8 fixed frequencies); out: [S][32][32][16].
int qnn_features_synth(const float* theta, const float* snr, int S, float recip, float semi_rad, float ang_rad,
                       float wavelength, const float* q_host, unsigned int seed, float* out, void* stream) {
  if (!theta || !snr || !q_host || !out || S < 1) return -1;
  FeatureOptics f{};
  f.recip = recip;
  f.semi = semi_rad;
  f.ang = ang_rad;
  f.lam = wavelength;
  for (int i = 0; i < 8; ++i) {
    f.qx[i] = q_host[2 * i];
    f.qy[i] = q_host[2 * i + 1];
  }
  hipLaunchKernelGGL((k_features_synth<8, 32>), dim3((unsigned)(S < 4096 ? S : 4096)), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), theta, snr, S, f, seed, out);
  return hipGetLastError() == hipSuccess ? 0 : -4;
}

}  // extern "C"
