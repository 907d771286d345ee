// Batched aberration search of the MI355X SSB engine: T trial aberration sets
// (C10, C12/phi12, C30 and the other even-order terms, each with its own scan
// rotation) evaluated on one or more datasets of possibly different scan size
// -- plans that hold a resident Hermitian Fourier stack -- in one persistent
// launch, replacing the reference's per-trial loop
// (src/quantem/gpu/ssb/compute/cuda/engine.py:1813-1836, driven by
// src/quantem/gpu/ssb/compute/cuda/optimizer.py:59-191).
//
// Objective: the variance over scan positions of the Fourier-sum object's
// phase, var_R angle(O_t(R)) with O_t = IFFT2(S_t), S_t(Q) = mean_b G_b(Q)
// conj(Gamma^_b(Q; theta_t)) and S_t(0) = dc (the object path of
// cuda/engine.py:1047-1107 for every trial).  It is a contrast measure:
// correct aberrations maximise it.  This is not the reference's default
// exact objective mean_R var_b phi_b (cuda/engine.py:1464-1514), which
// ssb_batch_phase_loss evaluates.
//
// Pipeline (all parameters read from device memory, so the sequence is
// captured once in a hipGraph and replayed for every new batch of trials):
//   k_trial_geometry   per (trial, BF pixel): rotated k, A(k), chi(k), Rout^2-|k|^2
//   k_trotter_trials   work queue over (dataset, block of TB trials, 256 points
//                      of the Hermitian half plane); each workgroup reads the
//                      stack G_b(Q) once per BF pixel for its TB trials
//   k_trials_cols      mirror S(-Q) = -conj S(Q) (even chi; the Nyquist row's
//                      other half is evaluated, not mirrored), /N_bf, DC, IFFT along q
//   k_trials_rows_var  IFFT along c, atan2, compensated sums of phi and phi^2
//   k_trials_var_final per trial: var = mean(phi^2) - mean(phi)^2
#include <cmath>
#include <string>
#include <vector>

#include "../../../include/quantem_ssb.h"
#include "ssb_internal.h"
#include "ssb_overlap.h"
#include "ssb_plan.h"

namespace qssb {

// Aberration phase and rotation of one trial, prepared on the host in double
// precision and rounded once to float32 (double-float for the rotation).
struct TrialChi {
  ChiPoly cp;       // even-m terms only (odd-m trials are rejected)
  int radial;       // only m = 0 terms: chi = r2 (K0 + r2 (K4 + r2 K10))
  int pad;
  df cos_a, sin_a;  // cos and sin of -rotation
};

// One (trial, BF pixel): rotated detector k as double-float, A(k), chi(k),
// and Rout^2 - |k|^2 as double-float.
struct TrialBf {
  float kx, kxl, ky, kyl;
  float ak, chik, cb, cbl;
};

// One dataset of the work queue.
struct TrialSet {
  const float2* Gt;      // resident Hermitian stack [nbf][N/2+1][N]
  const TrialBf* tbf;    // [T][nbf]
  float2* S;             // corrected Fourier sums [T][trial_points(N)]
  const float* qx;
  const float* qxl;
  const float* qy;
  const float* qyl;
  ApIso ap;
  int N, nbf, T, pad;
  long long item0;       // first queue item of this dataset
};

constexpr int kTrialBlock = 8;      // trials per workgroup of k_trotter_trials
// Corrected sums stored per trial: the Hermitian half plane (columns 0..N/2)
// and the Nyquist row's columns N/2+1..N-1.
__host__ __device__ constexpr int trial_points(int N) { return (N / 2 + 1) * N + (N / 2 - 1); }
constexpr int kTrialPoints = 256;   // half-plane points per workgroup

__device__ __forceinline__ float chi_even(const TrialChi& tc, float ux, float uy) {
  const float r2 = ux * ux + uy * uy;
  if (tc.radial) return r2 * (tc.cp.K[0] + r2 * (tc.cp.K[4] + r2 * tc.cp.K[10]));
  float e, o;
  chi_parts(tc.cp, ux, uy, e, o);
  return e;
}

__global__ void __launch_bounds__(256)
k_trial_geometry(const int* __restrict__ bf_rows, const int* __restrict__ bf_cols, const BfInfo* __restrict__ info,
                 int nbf, int T, df recip_y, df recip_x, Optics o, df rout2, const TrialChi* __restrict__ chis,
                 TrialBf* __restrict__ out) {
  const long long idx = (long long)blockIdx.x * 256 + threadIdx.x;
  if (idx >= (long long)T * nbf) return;
  const int t = (int)(idx / nbf), b = (int)(idx - (long long)t * nbf);
  const TrialChi& tc = chis[t];
  const df iy = two_sum((float)bf_rows[b], -info->center_r);
  const df ix = two_sum((float)bf_cols[b], -info->center_c);
  const df kxa = df_mul(iy, recip_y);
  const df kya = df_mul(ix, recip_x);
  const df kx = df_add(df_mul(kxa, tc.cos_a), df_mul(kya, tc.sin_a));
  const df ky = df_sub(df_mul(kya, tc.cos_a), df_mul(kxa, tc.sin_a));
  const df cb = df_sub(rout2, df_add(df_mul(kx, kx), df_mul(ky, ky)));
  TrialBf r;
  r.kx = kx.hi;
  r.kxl = kx.lo;
  r.ky = ky.hi;
  r.kyl = ky.lo;
  r.ak = aperture_band(kx, ky, o);
  r.chik = chi_even(tc, kx.hi, ky.hi);
  r.cb = cb.hi;
  r.cbl = cb.lo;
  out[idx] = r;
}

// Aperture from d = Rout^2 - |u|^2 given as a float32 estimate `dapprox`
// (absolute error far below kApMargin): exactly 1 or 0 outside the soft band
// plus the margin, otherwise -1 (evaluate d in double-float).
constexpr float kApMargin = 1e-4f;
__device__ __forceinline__ float aperture_quick(float dapprox, const ApIso& a) {
  if (dapprox >= a.d_in + kApMargin) return 1.0f;
  if (dapprox <= a.d_out - kApMargin) return 0.0f;
  return -1.0f;
}

// The persistent workgroups take items from a counter in device memory
// (queue[0]) until it passes total_items, so the long items (low columns,
// more points inside the apertures) and the short ones share the CUs evenly;
// the last workgroup to leave (queue[1]) sets both counters back to zero for
// the next launch, so a captured graph replays the kernel as it is.
// Two workgroups per CU (2 waves per SIMD, 256 registers with 2 spilled):
// unbounded, the kernel took 258 and ran one workgroup per CU at half the rate.
template <int TB>
__global__ void __launch_bounds__(256, 2)
k_trotter_trials(const TrialSet* __restrict__ sets, int n_sets, long long total_items,
                 const TrialChi* __restrict__ chis, unsigned* __restrict__ queue) {
  __shared__ TrialChi tch[TB];
  __shared__ unsigned s_item;
  const int tid = threadIdx.x;
  for (;;) {
    __syncthreads();   // the previous item's readers of tch and s_item are done
    if (tid == 0) s_item = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const long long item = s_item;
    if (item >= total_items) break;
    int d = 0;
    while (d + 1 < n_sets && item >= sets[d + 1].item0) ++d;
    const TrialSet ds = sets[d];
    const int N = ds.N, nbf = ds.nbf;
    const int NPG = (N / 2 + 1) * N;     // points of the Hermitian half plane (the stack's per-pixel size)
    const int NP = trial_points(N);      // plus the Nyquist row's other half
    const int pblocks = (NP + kTrialPoints - 1) / kTrialPoints;
    const long long loc = item - ds.item0;
    // point-block major: the low columns (the long items) are taken first
    const int tblocks = (ds.T + TB - 1) / TB;
    const int pb = (int)(loc / tblocks), tb = (int)(loc - (long long)pb * tblocks);
    const int t0 = tb * TB;
    const int nt = min(TB, ds.T - t0);
    {
      constexpr int W = (int)(sizeof(TrialChi) / 4);
      const int* src = reinterpret_cast<const int*>(chis + t0);
      int* dst = reinterpret_cast<int*>(tch);
      for (int w = tid; w < nt * W; w += 256) dst[w] = src[w];
    }
    __syncthreads();
    // point p: column c <= N/2, row q (the half plane, G read in place), or
    // p >= NPG: the Nyquist row q = N/2 at column c > N/2, whose mirror
    // -Q = (+N/2, N - c) is not a grid frequency (fftfreq puts -N/2 on the
    // grid), so it is evaluated directly with G(N/2, c) = conj G(N/2, N - c)
    const int p = pb * kTrialPoints + tid;
    const bool live = p < NP;
    const bool nyq = live && p >= NPG;
    const int c = nyq ? N / 2 + 1 + (p - NPG) : (live ? p / N : 0);
    const int q = nyq ? N / 2 : (live ? p - c * N : 0);
    const long long gofs = nyq ? (long long)(N - c) * N + N / 2 : (long long)(live ? p : 0);
    const df qx{ds.qx[q], ds.qxl[q]}, qy{ds.qy[c], ds.qyl[c]};
    const df q2 = df_add(df_mul(qx, qx), df_mul(qy, qy));
    const float qxf = qx.hi, qyf = qy.hi, q2f = q2.hi + q2.lo;
    Kahan ar[TB], ai[TB];
#pragma unroll
    for (int i = 0; i < TB; ++i) {
      ar[i].init();
      ai[i].init();
    }
    const float2* gsrc = ds.Gt + gofs;
    for (int b = 0; b < nbf; ++b) {
      float2 G = live ? gsrc[(long long)b * NPG] : make_float2(0.f, 0.f);
      if (nyq) G.y = -G.y;
#pragma unroll
      for (int i = 0; i < TB; ++i) {
        if (i < nt) {
          const TrialBf k = ds.tbf[(long long)(t0 + i) * nbf + b];
          const float qk = qxf * k.kx + qyf * k.ky;
          const float base = k.cb - q2f;
          float am = aperture_quick(base + 2.0f * qk, ds.ap);
          float ap = aperture_quick(base - 2.0f * qk, ds.ap);
          if (am < 0.0f || ap < 0.0f) {
            // inside the soft band: d = (Rout^2 - |k|^2 - |Q|^2) +/- 2 Q.k in double-float
            const df kx{k.kx, k.kxl}, ky{k.ky, k.kyl};
            const df qk2 = df_mul(df_add(df_mul(qx, kx), df_mul(qy, ky)), df{2.0f, 0.0f});
            const df bd = df_sub(df{k.cb, k.cbl}, q2);
            if (am < 0.0f) am = aperture_from_d(df_add(bd, qk2), ds.ap);
            if (ap < 0.0f) ap = aperture_from_d(df_sub(bd, qk2), ds.ap);
          }
          if (am != 0.0f || ap != 0.0f) {
            const float cm = chi_even(tch[i], qxf - k.kx, qyf - k.ky);
            const float cp = chi_even(tch[i], qxf + k.kx, qyf + k.ky);
            const float2 g = gamma_hat_chi(k.ak, am, ap, cm - k.chik, cp - k.chik);
            const float2 t = mul_conj(G, g);
            ar[i].add(t.x);
            ai[i].add(t.y);
          }
        }
      }
    }
    if (live) {
#pragma unroll
      for (int i = 0; i < TB; ++i)
        if (i < nt) ds.S[(long long)(t0 + i) * NP + p] = make_float2(ar[i].s - ar[i].c, ai[i].s - ai[i].c);
    }
  }
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&queue[0], 0u);
      atomicExch(&queue[1], 0u);
    }
  }
}

// Full corrected plane of trial blockIdx.y from its Hermitian half (even chi:
// S(-Q) = -conj S(Q)), divided by N_bf with S(0) = dc, inverse FFT along q
// of COUNT columns; W[t][c][q].
template <int N, int COUNT>
__global__ void __launch_bounds__(256)
k_trials_cols(const float2* __restrict__ S, const BfInfo* __restrict__ info, const float2* __restrict__ tw_g,
              float2* __restrict__ W) {
  constexpr int STRIDE = N + 1;
  constexpr int HALF = N / 2;
  constexpr long long NPG = (long long)(HALF + 1) * N;
  constexpr long long NP = trial_points(N);
  __shared__ float2 buf[COUNT * STRIDE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  const int t = blockIdx.y;
  const int c0 = blockIdx.x * COUNT;
  const float inv = 1.0f / (float)(info->nbf > 0 ? info->nbf : 1);
  const float2* st = S + (long long)t * NP;
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ci = idx / N, r = idx - (idx / N) * N;
    const int cc = c0 + ci;
    float2 v;
    if (cc <= HALF) {
      v = st[(long long)cc * N + r];
    } else if (r == HALF) {
      v = st[NPG + (cc - HALF - 1)];
    } else {
      const float2 m = st[(long long)(N - cc) * N + ((N - r) & (N - 1))];
      v = make_float2(-m.x, m.y);
    }
    v = (cc == 0 && r == 0) ? make_float2(info->dc_re, info->dc_im) : make_float2(v.x * inv, v.y * inv);
    buf[ci * STRIDE + r] = v;
  }
  __syncthreads();
  fft_lds<N, COUNT, 256, STRIDE, true>(buf, tw);
  float2* wt = W + (long long)t * N * N;
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ci = idx / N, r = idx - (idx / N) * N;
    wt[(long long)(c0 + ci) * N + r] = buf[ci * STRIDE + r];
  }
}

// compensated sum of (s, c) pairs over a 256-thread block, fixed order
__device__ void block_sum_pair(float& s, float& c, float* sh) {
#pragma clang fp contract(off)
  for (int off = 32; off > 0; off >>= 1) {
    const float os = __shfl_xor(s, off, 64);
    const float oc = __shfl_xor(c, off, 64);
    const df t = two_sum(s, os);
    s = t.hi;
    c = c + oc + t.lo;
  }
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) {
    sh[2 * w] = s;
    sh[2 * w + 1] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    float ts = sh[0], tc = sh[1];
    for (int k = 1; k < 4; ++k) {
      const df t = two_sum(ts, sh[2 * k]);
      ts = t.hi;
      tc = tc + sh[2 * k + 1] + t.lo;
    }
    s = ts;
    c = tc;
  }
}

// Inverse FFT along c of COUNT rows of trial blockIdx.y, phi = atan2, and
// the block's compensated sums of phi and phi^2 into part[t][block][4].
template <int N, int COUNT>
__global__ void __launch_bounds__(256)
k_trials_rows_var(const float2* __restrict__ W, const float2* __restrict__ tw_g, float* __restrict__ part) {
  constexpr int STRIDE = N + 1;
  __shared__ float2 buf[COUNT * STRIDE];
  __shared__ float2 tw[N];
  __shared__ float sh[8];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  const int t = blockIdx.y;
  const int r0 = blockIdx.x * COUNT;
  const float2* wt = W + (long long)t * N * N;
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int c = idx / COUNT, ri = idx - (idx / COUNT) * COUNT;
    buf[ri * STRIDE + c] = wt[(long long)c * N + r0 + ri];
  }
  __syncthreads();
  fft_lds<N, COUNT, 256, STRIDE, true>(buf, tw);
  Kahan k1, k2;
  k1.init();
  k2.init();
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ri = idx / N, c = idx - (idx / N) * N;
    const float2 v = buf[ri * STRIDE + c];
    const float ph = atan2f(v.y, v.x);
    k1.add(ph);
    k2.add(ph * ph);
  }
  float s1 = k1.s, c1 = -k1.c, s2 = k2.s, c2 = -k2.c;
  block_sum_pair(s1, c1, sh);
  block_sum_pair(s2, c2, sh);
  if (tid == 0) {
    float* o = part + ((long long)t * gridDim.x + blockIdx.x) * 4;
    o[0] = s1;
    o[1] = c1;
    o[2] = s2;
    o[3] = c2;
  }
}

// var_t = sum(phi^2)/N^2 - (sum(phi)/N^2)^2 from the blocks' sums, in block order.
__global__ void __launch_bounds__(64)
k_trials_var_final(const float* __restrict__ part, int blocks, int T, float inv_count, float* __restrict__ out) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= T) return;
  float s1 = 0.f, c1 = 0.f, s2 = 0.f, c2 = 0.f;
  for (int k = 0; k < blocks; ++k) {
    const float* o = part + ((long long)t * blocks + k) * 4;
    df a = two_sum(s1, o[0]);
    s1 = a.hi;
    c1 = c1 + o[1] + a.lo;
    df b = two_sum(s2, o[2]);
    s2 = b.hi;
    c2 = c2 + o[3] + b.lo;
  }
  const float m1 = (s1 + c1) * inv_count;
  const float m2 = (s2 + c2) * inv_count;
  out[t] = m2 - m1 * m1;
}

constexpr int kFinishCount = 8;   // columns / rows per finishing workgroup

hipError_t launch_trials_finish(int N, int T, const float2* S, const BfInfo* info, const float2* tw, float2* W,
                                float* part, float* out, hipStream_t s) {
  const dim3 grid(N / kFinishCount, T);
#define QSSB_FIN(NN)                                                                                      \
  if (N == NN) {                                                                                          \
    hipLaunchKernelGGL((k_trials_cols<NN, kFinishCount>), grid, dim3(256), 0, s, S, info, tw, W);         \
    hipLaunchKernelGGL((k_trials_rows_var<NN, kFinishCount>), grid, dim3(256), 0, s, W, tw, part);       \
  } else
  QSSB_FIN(128) QSSB_FIN(256) QSSB_FIN(512) QSSB_FIN(1024) return hipErrorInvalidValue;
#undef QSSB_FIN
  hipLaunchKernelGGL(k_trials_var_final, dim3((T + 63) / 64), dim3(64), 0, s, part, N / kFinishCount, T,
                     1.0f / ((float)N * (float)N), out);
  return hipGetLastError();
}

}  // namespace qssb

using namespace qssb;

// ---------------------------------------------------------------------------
// The batched-search object behind ssb_trials (quantem_ssb.h).
// ---------------------------------------------------------------------------
struct ssb_trials {
  std::vector<ssb_plan*> plans;
  int max_trials = 0;
  int n_trials = 0;
  std::vector<TrialChi> chis_host;
  TrialChi* chis = nullptr;
  TrialSet* sets = nullptr;
  std::vector<TrialSet> sets_host;
  std::vector<TrialBf*> tbf;
  std::vector<float2*> S;
  std::vector<float2*> W;
  std::vector<float*> part;
  int device = -1;
  hipStream_t stream = nullptr;          // the graph's own stream (never the legacy default)
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_trials = 0;                  // n_trials the graph was captured for
  float* graph_out = nullptr;
  unsigned* queue = nullptr;             // work-queue counters of k_trotter_trials (zero between launches)
};

namespace {

void trials_free(ssb_trials* tr) {
  if (!tr) return;
  if (tr->exec) (void)hipGraphExecDestroy(tr->exec);
  if (tr->graph) (void)hipGraphDestroy(tr->graph);
  if (tr->ev_in) (void)hipEventDestroy(tr->ev_in);
  if (tr->ev_out) (void)hipEventDestroy(tr->ev_out);
  if (tr->stream) (void)hipStreamDestroy(tr->stream);
  (void)hipFree(tr->chis);
  (void)hipFree(tr->queue);
  (void)hipFree(tr->sets);
  for (auto* p : tr->tbf) (void)hipFree(p);
  for (auto* p : tr->S) (void)hipFree(p);
  for (auto* p : tr->W) (void)hipFree(p);
  for (auto* p : tr->part) (void)hipFree(p);
  delete tr;
}

df to_df(double v) {
  const float hi = (float)v;
  return df{hi, (float)(v - (double)hi)};
}

// Enqueue the whole batched evaluation on `s`; out[p * max_trials + t].
hipError_t enqueue_trials(ssb_trials* tr, float* out, hipStream_t s) {
  const int T = tr->n_trials;
  for (size_t i = 0; i < tr->plans.size(); ++i) {
    ssb_plan* p = tr->plans[i];
    df rout2;
    (void)iso_aperture(p->optics, rout2);
    const long long total = (long long)T * p->g_nbf;
    hipLaunchKernelGGL(k_trial_geometry, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, p->bf_rows,
                       p->bf_cols, p->info, p->g_nbf, T, to_df(p->recip_y), to_df(p->recip_x), p->optics, rout2,
                       tr->chis, tr->tbf[i]);
  }
  long long total_items = 0;
  for (auto& ds : tr->sets_host) {
    const int NP = trial_points(ds.N);
    total_items += (long long)((T + kTrialBlock - 1) / kTrialBlock) * ((NP + kTrialPoints - 1) / kTrialPoints);
  }
  const long long grid = total_items < 2048 ? total_items : 2048;   // persistent: 8 per CU
  hipLaunchKernelGGL((k_trotter_trials<kTrialBlock>), dim3((unsigned)grid), dim3(256), 0, s, tr->sets,
                     (int)tr->sets_host.size(), total_items, tr->chis, tr->queue);
  for (size_t i = 0; i < tr->plans.size(); ++i) {
    ssb_plan* p = tr->plans[i];
    hipError_t e = launch_trials_finish(p->n, T, tr->S[i], p->info, p->tw, tr->W[i], tr->part[i],
                                        out + (long long)i * tr->max_trials, s);
    if (e != hipSuccess) return e;
  }
  return hipGetLastError();
}

}  // namespace

extern "C" {

int ssb_trials_create(ssb_plan* const* plans, int n_plans, int max_trials, ssb_trials** out) {
  clear_error();
  if (!out) return report_error(QSSB_ERR_ARG, "output pointer is NULL");
  *out = nullptr;
  if (!plans || n_plans < 1 || n_plans > 16) return report_error(QSSB_ERR_ARG, "between 1 and 16 plans are needed");
  if (max_trials < 1 || max_trials > 65536) return report_error(QSSB_ERR_ARG, "max_trials must be in [1, 65536]");
  for (int i = 0; i < n_plans; ++i) {
    ssb_plan* p = plans[i];
    if (!p) return report_error(QSSB_ERR_ARG, "plan is NULL");
    if (p->gcols != p->n / 2 + 1)
      return report_error(QSSB_ERR_STATE, "every plan needs a resident Hermitian Fourier stack (ssb_prepare_g_*)");
    if (!optics_isotropic(p->optics))
      return report_error(QSSB_ERR_UNSUPPORTED, "the batched search needs equal row and column angular sampling");
  }
  ssb_trials* tr = new ssb_trials();
  tr->plans.assign(plans, plans + n_plans);
  tr->max_trials = max_trials;
  hipError_t e = hipGetDevice(&tr->device);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&tr->chis), sizeof(TrialChi) * max_trials);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&tr->sets), sizeof(TrialSet) * n_plans);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&tr->queue), 2 * sizeof(unsigned));
  if (e == hipSuccess) e = hipMemset(tr->queue, 0, 2 * sizeof(unsigned));
  for (int i = 0; i < n_plans && e == hipSuccess; ++i) {
    ssb_plan* p = plans[i];
    const long long NP = trial_points(p->n);
    TrialBf* tb = nullptr;
    float2* S = nullptr;
    float2* W = nullptr;
    float* part = nullptr;
    e = hipMalloc(reinterpret_cast<void**>(&tb), sizeof(TrialBf) * (size_t)max_trials * p->g_nbf);
    if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&S), sizeof(float2) * (size_t)max_trials * NP);
    if (e == hipSuccess)
      e = hipMalloc(reinterpret_cast<void**>(&W), sizeof(float2) * (size_t)max_trials * p->n * p->n);
    if (e == hipSuccess)
      e = hipMalloc(reinterpret_cast<void**>(&part), sizeof(float) * 4 * (size_t)max_trials * (p->n / kFinishCount));
    tr->tbf.push_back(tb);
    tr->S.push_back(S);
    tr->W.push_back(W);
    tr->part.push_back(part);
    df rout2;
    TrialSet ds{};
    ds.Gt = p->Gt;
    ds.tbf = tb;
    ds.S = S;
    ds.qx = p->qx;
    ds.qxl = p->qxl;
    ds.qy = p->qy;
    ds.qyl = p->qyl;
    ds.ap = iso_aperture(p->optics, rout2);
    ds.N = p->n;
    ds.nbf = p->g_nbf;
    ds.T = 0;
    ds.item0 = 0;            // set with the trials (ssb_trials_set)
    tr->sets_host.push_back(ds);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&tr->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&tr->ev_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&tr->ev_out, hipEventDisableTiming);
  if (e != hipSuccess) {
    trials_free(tr);
    return report_hip_error(e, "ssb_trials_create");
  }
  *out = tr;
  return QSSB_OK;
}

int ssb_trials_destroy(ssb_trials* tr) {
  if (tr) (void)hipStreamSynchronize(tr->stream);
  trials_free(tr);
  return QSSB_OK;
}

int ssb_trials_set(ssb_trials* tr, int n_trials, const double* mags_host, const double* angles_host,
                   const double* rotation_host, void* stream) {
  clear_error();
  if (!tr) return report_error(QSSB_ERR_ARG, "trials handle is NULL");
  if (n_trials < 1 || n_trials > tr->max_trials)
    return report_error(QSSB_ERR_ARG, "n_trials must be in [1, " + std::to_string(tr->max_trials) + "]");
  if (!mags_host || !angles_host) return report_error(QSSB_ERR_ARG, "aberration arrays must not be NULL");
  static const int kn[14] = {1, 1, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5};
  static const int km[14] = {0, 2, 1, 3, 0, 2, 4, 1, 3, 5, 0, 2, 4, 6};
  const double lam = tr->plans[0]->wavelength;
  for (auto* p : tr->plans)
    if (p->wavelength != lam) return report_error(QSSB_ERR_ARG, "all plans must share the wavelength");
  tr->chis_host.assign(n_trials, TrialChi{});
  for (int t = 0; t < n_trials; ++t) {
    TrialChi& tc = tr->chis_host[t];
    const double* m = mags_host + (size_t)t * 14;
    const double* a = angles_host + (size_t)t * 14;
    tc.radial = 1;
    for (int i = 0; i < 14; ++i) {
      if ((km[i] & 1) && m[i] != 0.0)
        return report_error(QSSB_ERR_UNSUPPORTED, "the batched search evaluates even aberrations only "
                                                  "(C21, C23, C41, C43, C45 must be zero)");
      tc.cp.K[i] = (float)(2.0 * M_PI / lam * std::pow(lam, kn[i] + 1) / (kn[i] + 1) * m[i]);
      tc.cp.ca[i] = (float)std::cos(km[i] * a[i]);
      tc.cp.sa[i] = (float)std::sin(km[i] * a[i]);
      if (km[i] != 0 && tc.cp.K[i] != 0.0f) tc.radial = 0;
    }
    tc.cp.odd = 0;
    const double rot = rotation_host ? rotation_host[t] : 0.0;
    tc.cos_a = to_df(std::cos(-rot));
    tc.sin_a = to_df(std::sin(-rot));
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  // the work queue: each dataset's items follow the previous dataset's
  long long item0 = 0;
  for (auto& ds : tr->sets_host) {
    ds.T = n_trials;
    ds.item0 = item0;
    item0 += (long long)((n_trials + kTrialBlock - 1) / kTrialBlock) *
             ((trial_points(ds.N) + kTrialPoints - 1) / kTrialPoints);
  }
  hipError_t e = hipMemcpyAsync(tr->chis, tr->chis_host.data(), sizeof(TrialChi) * n_trials, hipMemcpyHostToDevice, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(tr->sets, tr->sets_host.data(), sizeof(TrialSet) * tr->sets_host.size(),
                       hipMemcpyHostToDevice, s);
  // the host arrays may change as soon as this returns
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return report_hip_error(e, "ssb_trials_set");
  tr->n_trials = n_trials;
  return QSSB_OK;
}

int ssb_trials_phase_variance(ssb_trials* tr, float* out, int use_graph, void* stream) {
  clear_error();
  if (!tr || !out) return report_error(QSSB_ERR_ARG, "trials handle and output must not be NULL");
  if (tr->n_trials < 1) return report_error(QSSB_ERR_STATE, "no trials set (ssb_trials_set)");
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return report_hip_error(e, "trials device");
  if (dev != tr->device) return report_error(QSSB_ERR_ARG, "the trials handle belongs to another device");
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  if (!use_graph) {
    e = enqueue_trials(tr, out, s);
    if (e != hipSuccess) return report_hip_error(e, "trials");
    return QSSB_OK;
  }
  if (!tr->exec || tr->graph_trials != tr->n_trials || tr->graph_out != out) {
    if (tr->exec) (void)hipGraphExecDestroy(tr->exec);
    if (tr->graph) (void)hipGraphDestroy(tr->graph);
    tr->exec = nullptr;
    tr->graph = nullptr;
    e = hipStreamBeginCapture(tr->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return report_hip_error(e, "trials capture");
    hipError_t ee = enqueue_trials(tr, out, tr->stream);
    e = hipStreamEndCapture(tr->stream, &tr->graph);
    if (ee != hipSuccess) return report_hip_error(ee, "trials capture enqueue");
    if (e != hipSuccess) return report_hip_error(e, "trials end capture");
    e = hipGraphInstantiate(&tr->exec, tr->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return report_hip_error(e, "trials graph instantiate");
    tr->graph_trials = tr->n_trials;
    tr->graph_out = out;
  }
  // the graph runs on the handle's own stream, ordered after and before `s`
  e = hipEventRecord(tr->ev_in, s);
  if (e == hipSuccess) e = hipStreamWaitEvent(tr->stream, tr->ev_in, 0);
  if (e == hipSuccess) e = hipGraphLaunch(tr->exec, tr->stream);
  if (e == hipSuccess) e = hipEventRecord(tr->ev_out, tr->stream);
  if (e == hipSuccess) e = hipStreamWaitEvent(s, tr->ev_out, 0);
  if (e != hipSuccess) return report_hip_error(e, "trials graph launch");
  return QSSB_OK;
}

}  // extern "C"
