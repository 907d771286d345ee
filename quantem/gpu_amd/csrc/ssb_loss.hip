// Exact per-BF phase path of the MI355X SSB engine: for every bright-field
// pixel b the corrected plane C_b(Q) = G_b(Q) conj(Gamma^_b(Q)) with its DC
// bin replaced by dc, the inverse 2D FFT O_b = IFFT2(C_b), phi_b = angle(O_b),
// and the per-pixel moments sum_b phi_b, sum_b phi_b^2.  From them the mean
// phase image phase = mean_b phi_b and the phase-variance loss
// L = mean_R(mean_b phi_b^2 - phase^2).
//
// Reference behaviour (paths relative to bobleesj/quantem.gpu):
//   reconstruct / reconstruct_with_loss  src/quantem/gpu/ssb/compute/cuda/engine.py:1153-1191, 1464-1514
//   fused gamma + IFFT                   src/quantem/gpu/ssb/compute/cuda/kernels/common.py:768-816
//   phase moments                        src/quantem/gpu/ssb/compute/cuda/engine.py:21-73
//   chunked moments                      src/quantem/gpu/ssb/compute/cuda/engine.py:1547-1783
//
// Two passes over a chunk of BF pixels, with the chunk's column-transformed
// planes T[b][x][c] (complex64, 64 MB, and 128 MB at scan 256 where the
// larger chunk measured 3-5% faster per trial; both stay in the 256 MB
// last-level cache between the passes):
//   1. k_corr_colifft: per (b, Hermitian slot group) build the corrected
//      columns C_b(:, j) and C_b(:, N-j) from the resident half-plane stack
//      (one aperture pair and one phasor serve both), inverse FFT along q in
//      LDS, write T;
//   2. k_rowifft_moments: per (row block, BF split) inverse FFT along c,
//      atan2, Kahan-compensated moments carried across chunks in
//      (sum, compensation) planes M1/M2[split][x][c].
// A final pass combines the splits in fixed order (bit-repeatable).
#include "ssb_internal.h"
#include "ssb_overlap.h"

namespace qssb {

// ------------------------------------------------------------------------
// Pass 1.  Work item = (b in chunk, slot group sg).  Output columns: slot s
// of the group gives A = column j = sg*SG + s and B = column N - j (rows
// mirrored, from conj G) or, for j == 0, column N/2.  BLOCK is a multiple
// of N; each thread owns one row q.  FULL: the resident stack holds all N
// columns (an imported full-plane stack), so the B column is read, not
// mirrored from A.
// ------------------------------------------------------------------------
template <int N, int SG, int BLOCK, bool ISO, bool FULL, bool POLY>
__global__ void __launch_bounds__(BLOCK)
k_corr_colifft(const float2* __restrict__ Gt, const BfInfo* __restrict__ info, GeomPtrs gp,
               const float2* __restrict__ rowph_g, const float2* __restrict__ colph_g,
               const float2* __restrict__ echi_g, ChiPoly cp_arg, const ChiPoly* __restrict__ cp_dev, ApIso ap,
               df rout2, Optics o, const float2* __restrict__ tw_g, int b_begin, int b_count,
               float2* __restrict__ T) {
  // the chi polynomial: a kernel argument, or read from device memory (a
  // captured batch of trials replays with new coefficients)
  const ChiPoly cp = (POLY && cp_dev) ? *cp_dev : cp_arg;
  constexpr int HALF = N / 2;
  constexpr int NSG = HALF / SG;
  constexpr int P = SG * N / BLOCK;
  constexpr int SSTEP = BLOCK / N;
  constexpr int GCOLS = HALF + 1;           // columns of the phasor table
  constexpr int SCOLS = FULL ? N : GCOLS;   // columns of the stack
  constexpr int STRIDE = N + 1;
  constexpr int NC = 2 * SG;
  static_assert(BLOCK % N == 0 && P >= 1, "block must cover whole rows");
  __shared__ float2 buf[NC * STRIDE];
  __shared__ float2 tw[N];
  __shared__ float qyS[SG + 1], qySl[SG + 1];
  __shared__ float2 ehalfc[N];
  __shared__ df q2halfc[N];
  __shared__ float2 enyqr[SG];
  // ISO column terms of d(Q -/+ k): (Rout^2 - |k|^2) - qy^2 +/- 2 qy ky
  __shared__ df colA_s[SG + 1], colB_s[SG + 1];

  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BLOCK) tw[i] = tw_g[i];
  const int nbf = info->nbf;
  const float2 dcv = make_float2(info->dc_re, info->dc_im);
  int b_end = b_begin + b_count;
  if (b_end > nbf) b_end = nbf;
  const long long items = (b_end > b_begin) ? (long long)(b_end - b_begin) * NSG : 0;
  const int q = tid % N;
  const int s0 = tid / N;
  const int qm = (N - q) & (N - 1);
  const df qxA{gp.qx[q], gp.qxl[q]};
  const df qxB0{gp.qx[qm], gp.qxl[qm]};
  const df q2x = df_mul(qxA, qxA);

  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    const int bl = (int)(item / NSG);
    const int sg = (int)(item - (long long)bl * NSG);
    const int b = b_begin + bl;
    __syncthreads();   // the previous item's readers of buf and the tables are done
    if (tid <= SG) {
      const int col = (tid < SG) ? sg * SG + tid : HALF;
      qyS[tid] = gp.qy[col];
      qySl[tid] = gp.qyl[col];
    }
    if (sg == 0 && tid < N) {
      const df qx{gp.qx[tid], gp.qxl[tid]};
      const df qh{gp.qy[HALF], gp.qyl[HALF]};
      if (!ISO) q2halfc[tid] = df_add(df_mul(qx, qx), df_mul(qh, qh));
      if (!POLY) ehalfc[tid] = echi_g[(long long)HALF * N + tid];
    }
    if (!POLY && tid < SG) {
      const int j = sg * SG + tid;
      enyqr[tid] = echi_g[(long long)((N - j) & (N - 1)) * N + HALF];
    }
    const df kx{gp.kx[b], gp.kxl[b]}, ky{gp.ky[b], gp.kyl[b]};
    if (ISO && tid <= SG) {
      const int col = (tid < SG) ? sg * SG + tid : HALF;
      const df qy{gp.qy[col], gp.qyl[col]};
      const df cbv = df_sub(rout2, df_add(df_mul(kx, kx), df_mul(ky, ky)));
      const df base = df_sub(cbv, df_mul(qy, qy));
      const df qk = df_mul(qy, df{2.0f * ky.hi, 2.0f * ky.lo});
      colA_s[tid] = df_add(base, qk);
      colB_s[tid] = df_sub(base, qk);
    }
    __syncthreads();
    const float ak = gp.ak[b];
    const float2 rowA = POLY ? make_float2(1.f, 0.f) : rowph_g[(long long)b * N + q];
    const float2 colHalf = POLY ? make_float2(1.f, 0.f) : colph_g[(long long)b * GCOLS + HALF];
    float chik = 0.0f;
    if constexpr (POLY) {
      float e, od;
      chi_parts(cp, kx.hi, ky.hi, e, od);
      chik = e + od;
    }
    // ISO row terms of d(Q -/+ k) as float32 estimates (double-float only
    // for points near the soft edge)
    const df kx2{2.0f * kx.hi, 2.0f * kx.lo};
    float uq = 0.0f, vq = 0.0f;
    df cb{0.f, 0.f};
    if constexpr (ISO) {
      uq = row_estimate_minus(qxA, kx2, q2x);
      vq = row_estimate_plus(qxA, kx2, q2x);
    } else {
      cb = df_sub(rout2, df_add(df_mul(kx, kx), df_mul(ky, ky)));
    }
    const float2* gb = Gt + (long long)b * SCOLS * N;
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int s = s0 + i * SSTEP;
      const int j = sg * SG + s;
      const float2 z = gb[(long long)j * N + q];
      const float2 GA = z;
      const float2 GB = (j == 0) ? gb[(long long)HALF * N + q]
                                 : (FULL ? gb[(long long)(N - j) * N + qm] : cconj(z));
      const df qy{qyS[s], qySl[s]};
      const df q2A = ISO ? df{0.f, 0.f} : df_add(q2x, df_mul(qy, qy));
      const float2 echiA = POLY ? make_float2(1.f, 0.f) : echi_g[(long long)j * N + q];
      const float2 colp = POLY ? make_float2(1.f, 0.f) : colph_g[(long long)b * GCOLS + j];
      float am, app;
      if constexpr (ISO) {
        iso_aperture_pair(uq, vq, qxA, kx2, q2x, colA_s[s], colB_s[s], ap, am, app);
      } else {
        aperture_pair<ISO>(qxA, qy, kx, ky, cb, q2A, ap, o, am, app);
      }
      const float2 eA = cmul(rowA, colp);
      const bool mirror = (j != 0) && (q != HALF);
      float bm, bp;
      float2 eB, echiB;
      if (mirror) {
        bm = app;
        bp = am;
        eB = cconj(eA);
        echiB = echiA;
      } else {
        echiB = POLY ? make_float2(1.f, 0.f) : ((j == 0) ? ehalfc[q] : enyqr[s]);
        if constexpr (ISO) {
          iso_aperture_pair(uq, vq, qxA, kx2, q2x, (j == 0) ? colA_s[SG] : colB_s[s],
                            (j == 0) ? colB_s[SG] : colA_s[s], ap, bm, bp);
        } else {
          const df qyB = (j == 0) ? df{qyS[SG], qySl[SG]} : df{-qy.hi, -qy.lo};
          const df qxB = (j == 0) ? qxA : qxB0;
          const df q2B = (j == 0) ? q2halfc[q] : q2A;
          aperture_pair<ISO>(qxB, qyB, kx, ky, cb, q2B, ap, o, bm, bp);
        }
        eB = cmul(rowA, (j == 0) ? colHalf : cconj(colp));
      }
      // Mirror points with an even chi (every quadratic chi; a full chi
      // without odd-m terms): B sees the apertures swapped, the phasors
      // conjugated and conj(G), so Gamma_B = -conj(Gamma_A) with the same
      // norm and C_B = -conj(C_A) exactly (every operation is
      // sign-symmetric); it is not recomputed.
      const bool same = mirror && !FULL && (!POLY || !cp.odd);
      float2 gA, gB = make_float2(0.f, 0.f);
      if constexpr (POLY) {
        // chi at Q -/+ k split into even and odd parts: chi(-u) = E(u) - O(u)
        const float qxf = qxA.hi, qyf = qy.hi;
        float em, om, ep, op;
        chi_parts(cp, qxf - kx.hi, qyf - ky.hi, em, om);
        chi_parts(cp, qxf + kx.hi, qyf + ky.hi, ep, op);
        gA = gamma_hat_chi(ak, am, app, em + om - chik, ep + op - chik);
        if (same) {
        } else if (mirror) {
          gB = gamma_hat_chi(ak, bm, bp, ep - op - chik, em - om - chik);
        } else {
          const float qxb = (j == 0) ? qxf : qxB0.hi;
          const float qyb = (j == 0) ? qyS[SG] : -qyf;
          float e1, o1, e2, o2;
          chi_parts(cp, qxb - kx.hi, qyb - ky.hi, e1, o1);
          chi_parts(cp, qxb + kx.hi, qyb + ky.hi, e2, o2);
          gB = gamma_hat_chi(ak, bm, bp, e1 + o1 - chik, e2 + o2 - chik);
        }
      } else {
        gA = gamma_hat(ak, eA, am, app, echiA);
        if (!same) gB = gamma_hat(ak, eB, bm, bp, echiB);
      }
      float2 CA = mul_conj(GA, gA);
      const float2 CB = same ? make_float2(-CA.x, CA.y) : mul_conj(GB, gB);
      if (j == 0 && q == 0) CA = dcv;
      const int rowB = (j == 0) ? q : qm;
      buf[s * STRIDE + q] = CA;
      buf[(SG + s) * STRIDE + rowB] = CB;
    }
    __syncthreads();
    fft_lds<N, NC, BLOCK, STRIDE, true>(buf, tw);
    float2* tb = T + (long long)bl * N * N;
    for (int idx = tid; idx < NC * N; idx += BLOCK) {
      const int x = idx / NC;
      const int ci = idx - x * NC;
      const int s = (ci < SG) ? ci : ci - SG;
      const int j = sg * SG + s;
      const int col = (ci < SG) ? j : ((j == 0) ? HALF : N - j);
      tb[(long long)x * N + col] = buf[ci * STRIDE + x];
    }
  }
}

// ------------------------------------------------------------------------
// Pass 2.  Block = (ROWS rows x0.., BF split).  The split's share of the
// chunk is [split*cnt/nsplit, (split+1)*cnt/nsplit).  Moments are carried in
// M1/M2[split][x][c] as (Kahan sum, compensation); `first` starts them at 0.
// ------------------------------------------------------------------------
template <int N, int ROWS, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_rowifft_moments(const float2* __restrict__ T, const BfInfo* __restrict__ info, int b_begin, int b_count,
                  int nsplit, const float2* __restrict__ tw_g, float2* __restrict__ M1, float2* __restrict__ M2,
                  int first) {
  constexpr int STRIDE = N + 1;
  constexpr int E = ROWS * N / BLOCK;
  static_assert(E >= 1 && (ROWS * N) % BLOCK == 0, "block must tile the rows");
  __shared__ float2 buf[ROWS * STRIDE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BLOCK) tw[i] = tw_g[i];
  const int x0 = blockIdx.x * ROWS;
  const int split = blockIdx.y;
  int cnt = info->nbf - b_begin;
  if (cnt > b_count) cnt = b_count;
  if (cnt < 0) cnt = 0;
  const int b0 = (int)(((long long)split * cnt) / nsplit);
  const int b1 = (int)(((long long)(split + 1) * cnt) / nsplit);
  const long long NN = (long long)N * N;
  float2* m1 = M1 + (long long)split * NN;
  float2* m2 = M2 + (long long)split * NN;
  Kahan k1[E], k2[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = tid + e * BLOCK;
    const long long o = (long long)(x0 + idx / N) * N + (idx % N);
    if (first) {
      k1[e].init();
      k2[e].init();
    } else {
      const float2 a = m1[o], c = m2[o];
      k1[e].s = a.x;
      k1[e].c = a.y;
      k2[e].s = c.x;
      k2[e].c = c.y;
    }
  }
  // the next BF pixel's rows are loaded into registers while this one's
  // are transformed, so the loads' latency overlaps the FFT
  float2 nxt[E];
  auto load = [&](int b) {
    const float2* tb = T + (long long)b * NN + (long long)x0 * N;
#pragma unroll
    for (int e = 0; e < E; ++e) nxt[e] = tb[tid + e * BLOCK];
  };
  if (b0 < b1) load(b0);
  for (int b = b0; b < b1; ++b) {
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int idx = tid + e * BLOCK;
      buf[(idx / N) * STRIDE + (idx % N)] = nxt[e];
    }
    if (b + 1 < b1) load(b + 1);
    __syncthreads();
    fft_lds<N, ROWS, BLOCK, STRIDE, true>(buf, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int idx = tid + e * BLOCK;
      const float2 v = buf[(idx / N) * STRIDE + (idx % N)];
      const float ph = atan2f(v.y, v.x);
      k1[e].add(ph);
      k2[e].add(ph * ph);
    }
  }
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int idx = tid + e * BLOCK;
    const long long o = (long long)(x0 + idx / N) * N + (idx % N);
    m1[o] = make_float2(k1[e].s, k1[e].c);
    m2[o] = make_float2(k2[e].s, k2[e].c);
  }
}

// ------------------------------------------------------------------------
// Pass 3: combine the splits (fixed order), mean phase, per-pixel variance,
// per-block variance sums; then one block sums those in fixed order.
// ------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_moments_finish(const float2* __restrict__ M1, const float2* __restrict__ M2, int nsplit, int NN,
                 const BfInfo* __restrict__ info, float* __restrict__ phase, float* __restrict__ var_part) {
  __shared__ float red_s[4];
  const int idx = blockIdx.x * 256 + threadIdx.x;
  float var = 0.0f;
  if (idx < NN) {
    Kahan s1, s2;
    s1.init();
    s2.init();
    for (int sp = 0; sp < nsplit; ++sp) {
      const float2 a = M1[(long long)sp * NN + idx];
      const float2 c = M2[(long long)sp * NN + idx];
      s1.add(a.x - a.y);
      s2.add(c.x - c.y);
    }
    const float inv = 1.0f / (float)(info->nbf > 0 ? info->nbf : 1);
    const float mean = (s1.s - s1.c) * inv;
    var = (s2.s - s2.c) * inv - mean * mean;
    if (phase) phase[idx] = mean;
  }
  // block sum of var: wave tree sums (fixed order), then the 4 waves in order
  float v = var;
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red_s[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    Kahan k;
    k.init();
    for (int i = 0; i < 4; ++i) k.add(red_s[i]);
    var_part[blockIdx.x] = k.s - k.c;
  }
}

__global__ void __launch_bounds__(256)
k_loss_final(const float* __restrict__ var_part, int nparts, int NN, float* __restrict__ loss) {
  __shared__ float ws[4], wc[4];
  Kahan k;
  k.init();
  for (int i = threadIdx.x; i < nparts; i += 256) k.add(var_part[i]);
  float s = k.s, c = k.c;
  // combine lanes in a fixed tree; carry compensation separately
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_down(s, off, 64);
    c += __shfl_down(c, off, 64);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    ws[w] = s;
    wc[w] = c;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Kahan t;
    t.init();
    for (int i = 0; i < 4; ++i) t.add(ws[i] - wc[i]);
    loss[0] = (t.s - t.c) / (float)NN;
  }
}

// ------------------------------------------------------------------------
// Launch wrappers
// ------------------------------------------------------------------------
int loss_chunk_bf(int N) { return N == 128 ? 512 : (N == 256 ? 256 : (N == 512 ? 32 : 8)); }
// BF splits of the row pass (partial moment planes): 16 at N = 256 keeps
// 2048 workgroups in flight (8 measured 2.6% slower per trial)
int loss_splits(int N) { return N == 128 ? 32 : (N == 256 ? 16 : (N == 512 ? 2 : 1)); }

hipError_t launch_phase_moments(int N, const float2* Gt, int gcols, const BfInfo* info, int nbf_hint, const GeomPtrs& gp,
                                const float2* rowph, const float2* colph, const float2* echi, const ChiPoly* poly,
                                const Optics& o,
                                const float2* tw, float2* T, float2* M1, float2* M2, float* var_part,
                                float* phase, float* loss, hipStream_t s, const ChiPoly* poly_dev) {
  df rout2;
  const ApIso ap = iso_aperture(o, rout2);
  const bool iso = optics_isotropic(o);
  const bool full = (gcols == N);
  if (!full && gcols != N / 2 + 1) return hipErrorInvalidValue;
  const int chunk = loss_chunk_bf(N);
  const int nsplit = loss_splits(N);
  const int nb = nbf_hint > 0 ? nbf_hint : 1;
#define QSSB_P1_ONE(NN, SGV, BL, IS, FU)                                                                    \
  do {                                                                                                       \
    if (poly)                                                                                                \
      hipLaunchKernelGGL((k_corr_colifft<NN, SGV, BL, IS, FU, true>), dim3(g1), dim3(BL), 0, s, Gt, info, gp,   \
                         rowph, colph, echi, *poly, poly_dev, ap, rout2, o, tw, b0, cnt, T);                  \
    else                                                                                                     \
      hipLaunchKernelGGL((k_corr_colifft<NN, SGV, BL, IS, FU, false>), dim3(g1), dim3(BL), 0, s, Gt, info, gp,  \
                         rowph, colph, echi, ChiPoly{}, nullptr, ap, rout2, o, tw, b0, cnt, T);               \
  } while (0)
#define QSSB_P1(NN, SGV, BL)                                                    \
  do {                                                                         \
    const unsigned g1 = (unsigned)((long long)cnt * ((NN / 2) / SGV));        \
    if (iso && !full) QSSB_P1_ONE(NN, SGV, BL, true, false);                   \
    else if (iso) QSSB_P1_ONE(NN, SGV, BL, true, true);                        \
    else if (!full) QSSB_P1_ONE(NN, SGV, BL, false, false);                    \
    else QSSB_P1_ONE(NN, SGV, BL, false, true);                                \
  } while (0)
#define QSSB_P2(NN, ROWS)                                                                                    \
  hipLaunchKernelGGL((k_rowifft_moments<NN, ROWS, 256>), dim3(NN / ROWS, nsplit), dim3(256), 0, s, T, info, b0, \
                     cnt, nsplit, tw, M1, M2, b0 == 0 ? 1 : 0)
  for (int b0 = 0; b0 < nb; b0 += chunk) {
    const int cnt = (nb - b0 < chunk) ? nb - b0 : chunk;
    switch (N) {
      case 128:
        QSSB_P1(128, 8, 512);
        QSSB_P2(128, 8);
        break;
      case 256:
        // quadratic chi: 16 slot columns per 1024-thread item (6% faster);
        // full chi: 8 per 512 (the wider item is no faster there)
        if (poly) QSSB_P1(256, 8, 512);
        else QSSB_P1(256, 16, 1024);
        QSSB_P2(256, 2);
        break;
      case 512:
        QSSB_P1(512, 4, 512);
        QSSB_P2(512, 2);
        break;
      case 1024:
        QSSB_P1(1024, 2, 1024);
        QSSB_P2(1024, 1);
        break;
      default:
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
#undef QSSB_P1
#undef QSSB_P1_ONE
#undef QSSB_P2
  const int NN = N * N;
  const int parts = (NN + 255) / 256;
  hipLaunchKernelGGL(k_moments_finish, dim3(parts), dim3(256), 0, s, M1, M2, nsplit, NN, info, phase, var_part);
  hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(256), 0, s, var_part, parts, NN, loss);
  return hipGetLastError();
}

}  // namespace qssb
