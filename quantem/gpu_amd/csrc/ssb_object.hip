// Object-path kernels of the MI355X SSB engine: the trotter reduction over
// bright-field pixels (fused with the scan-column FFT when it starts from the
// row-transformed stack H, or reading a resident Fourier stack), and the
// final inverse FFT to the object and its phase.
//
// Reference behaviour (paths relative to bobleesj/quantem.gpu):
//   Fourier-sum trotter   src/quantem/gpu/ssb/compute/cuda/kernels/fft256.py:967-1030
//   overlap / guard       src/quantem/gpu/ssb/compute/cuda/kernels/common.py:115-236
//   object + IFFT         src/quantem/gpu/ssb/compute/cuda/engine.py:1047-1107
//   phase = angle(object) src/quantem/gpu/ssb/results.py:117-122
#include "ssb_internal.h"
#include "ssb_overlap.h"

namespace qssb {

int object_groups(int N) { return N == 128 ? 64 : (N == 256 ? 16 : (N == 512 ? 8 : 4)); }

// The phasor tables (PhasorJob) on their own launch; m_dev, when given,
// holds the quadratic chi instead of m_arg.
__global__ void __launch_bounds__(256)
k_phasors(const BfInfo* __restrict__ info, GeomPtrs gp, QuadChi m_arg, int N, float2* __restrict__ rowph,
          float2* __restrict__ colph, float2* __restrict__ echi, const QuadChi* __restrict__ m_dev) {
  const PhasorJob j{info, gp, m_dev ? *m_dev : m_arg, N, rowph, colph, echi};
  phasor_work(j, (long long)blockIdx.x * 256 + threadIdx.x, (long long)gridDim.x * 256);
}

// ------------------------------------------------------------------------
// Trotter over Hermitian slot columns.  Work item = (BF group g, slot group
// sg of SG slots).  Per BF pixel the slot columns G_b(:, j) are staged in
// LDS -- from H through the scan-column FFT (FROM_H) or copied from the
// resident stack Gt[b][c][q] -- and each G_b(q, j) contributes to
//   output A = S(q, j)          with G_b(q, j), and
//   output B = S(-q, N-j)       with conj(G_b(q, j))        (j >= 1),
// or, for slot 0, A = S(q, 0) with G_b(q, 0) and B = S(q, N/2) with
// G_b(q, N/2) (slot 0 of H packs the two real columns 0 and N/2).
// For j >= 1 and q != N/2, Q_B = -Q_A, so chi(Q_B) = chi(Q_A),
// theta_B = -theta_A, A(Q_B -/+ k) = A(Q_A +/- k): one aperture pair and one
// phasor serve both outputs.  Points whose two apertures vanish contribute
// exactly zero and are skipped.
// BLOCK is a multiple of N: each thread keeps one scan-frequency row q.
// ------------------------------------------------------------------------
template <int N, int SG, int BLOCK, bool FROM_H, bool ISO>
__global__ void __launch_bounds__(BLOCK)
k_trotter_slots(const float2* __restrict__ src, const BfInfo* __restrict__ info, GeomPtrs gp,
                const float2* __restrict__ rowph_g, const float2* __restrict__ colph_g,
                const float2* __restrict__ echi_g, ApIso ap, df rout2, Optics o,
                const float2* __restrict__ tw_g, int groups, float2* __restrict__ S_part, float2* __restrict__ Gt,
                int do_trotter) {
  constexpr int HALF = N / 2;
  constexpr int NSG = HALF / SG;
  constexpr int P = SG * N / BLOCK;      // points per thread
  constexpr int SSTEP = BLOCK / N;       // slot step between a thread's points
  constexpr int GCOLS = HALF + 1;
  constexpr int TILE = N * SG;           // complex values per BF tile
  constexpr int COPIES = (TILE + 2 * BLOCK - 1) / (2 * BLOCK);
  constexpr int TSTRIDE = N + 1;
  // slot groups sharing one 128-byte line of H
  constexpr int LG = (FROM_H && SG * 8 < 128) ? 128 / (SG * 8) : 1;
  static_assert(TILE % 128 == 0, "tile must split into whole waves of 16-byte copies");
  // a lane's 16-byte copy takes two adjacent slots of one row of H
  static_assert(!FROM_H || SG % 2 == 0, "row-transformed stacks need an even slot count");
  // Async landing buffers (double-buffered).  FROM_H: interleaved rows of H,
  // element (s, r) at [r * SG + s]; otherwise Gt columns, (s, q) at [s * N + q].
  __shared__ __attribute__((aligned(16))) float2 abuf[2][TILE];
  // FROM_H: transformed columns, (s, q) at [s * TSTRIDE + q]
  __shared__ float2 tileT[FROM_H ? SG * TSTRIDE : 1];
  __shared__ float2 tw[FROM_H ? N : 1];
  __shared__ float qyS[SG + 1], qySl[SG + 1];
  // output-B constants where Q_B != -Q_A: column N/2 (slot 0) per row, and
  // the Nyquist row q = N/2 per slot
  __shared__ float2 ehalfc[N];
  __shared__ df q2halfc[N];
  __shared__ float2 enyqr[SG];
  // ISO: column terms of d(Q -/+ k) for the SG slot columns and column N/2,
  // double-buffered over BF pixels:
  //   colA = (Rout^2 - |k|^2) - qy^2 + 2 qy ky,  colB = (Rout^2 - |k|^2) - qy^2 - 2 qy ky
  __shared__ df colA_s[2][SG + 1], colB_s[2][SG + 1];

  const int tid = threadIdx.x;
  if (FROM_H)
    for (int i = tid; i < N; i += BLOCK) tw[i] = tw_g[i];
  const int nbf = info->nbf;
  const long long items = (long long)groups * NSG;
  const int q = tid % N;
  const int s0 = tid / N;
  const int qm = (N - q) & (N - 1);
  const df qxA{gp.qx[q], gp.qxl[q]};
  const df qxB0{gp.qx[qm], gp.qxl[qm]};   // row of output B when j >= 1
  const df q2x = df_mul(qxA, qxA);
  const int wave_base = tid & ~63;

  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    // XCD-aware order: workgroups are dispatched round-robin over the 8 XCDs,
    // so items i, i+8, ..., i+8(LG-1) run together on one XCD; give them the
    // LG slot groups that share each 128-byte line of H (read once per L2).
    // From a resident stack, slot groups in increasing order over the BF
    // groups: the low slot columns (more points inside the apertures, the
    // long items) start first (measured faster there; from H the BF-group
    // order is kept: the other order measured slower at scan 128).
    const long long sl = (item >> 3) % LG;
    const long long t = (item / (8 * LG)) * 8 + (item & 7);
    const int g = FROM_H ? (int)(t / (NSG / LG)) : (int)(t % groups);
    const int sg = FROM_H ? (int)((t % (NSG / LG)) * LG + sl) : (int)((t / groups) * LG + sl);
    int b0, b1;
    group_range(g, groups, nbf, b0, b1);
    // copy the source tile of BF pixel b into landing buffer `slot`
    auto issue = [&](int b, int slot) {
#pragma unroll
      for (int c = 0; c < COPIES; ++c) {
        const int e0 = (c * BLOCK + wave_base) * 2;         // wave's first element
        if (e0 >= TILE) continue;                           // wave-uniform
        const int el = e0 + (tid & 63) * 2;                 // this lane's first element
        const float2* gsrc;
        if (FROM_H) {
          const int r = el / SG, sl = el % SG;
          gsrc = src + ((long long)b * N + r) * HALF + sg * SG + sl;
        } else {
          gsrc = src + ((long long)b * GCOLS + sg * SG) * N + el;
        }
        copy_async16(gsrc, &abuf[slot][e0]);
      }
    };
    // column terms of BF pixel b into buffer `slot` (threads 0..SG)
    auto colterms = [&](int b, int slot) {
      if (ISO && tid <= SG) {
        const int col = (tid < SG) ? sg * SG + tid : HALF;
        const df qy{gp.qy[col], gp.qyl[col]};
        const df kx{gp.kx[b], gp.kxl[b]}, ky{gp.ky[b], gp.kyl[b]};
        const df cbv = df_sub(rout2, df_add(df_mul(kx, kx), df_mul(ky, ky)));
        const df base = df_sub(cbv, df_mul(qy, qy));
        const df qk = df_mul(qy, df{2.0f * ky.hi, 2.0f * ky.lo});
        colA_s[slot][tid] = df_add(base, qk);
        colB_s[slot][tid] = df_sub(base, qk);
      }
    };
    __syncthreads();
    if (b0 < b1) issue(b0, 0);
    if (b0 < b1) colterms(b0, 0);
    if (tid <= SG) {
      const int col = (tid < SG) ? sg * SG + tid : HALF;
      qyS[tid] = gp.qy[col];
      qySl[tid] = gp.qyl[col];
    }
    if (do_trotter) {
      if (sg == 0 && tid < N) {
        const df qx{gp.qx[tid], gp.qxl[tid]};
        const df qh{gp.qy[HALF], gp.qyl[HALF]};   // not qyS: written in this same phase
        if (!ISO) q2halfc[tid] = df_add(df_mul(qx, qx), df_mul(qh, qh));
        ehalfc[tid] = echi_g[(long long)HALF * N + tid];
      }
      if (tid < SG) {
        const int j = sg * SG + tid;
        enyqr[tid] = echi_g[(long long)((N - j) & (N - 1)) * N + HALF];
      }
    }
    __syncthreads();
    // per-point constants
    float2 echiA[P];
    df q2A[P];
    Kahan arA[P], aiA[P], arB[P], aiB[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int s = s0 + i * SSTEP;
      const df qy{qyS[s], qySl[s]};
      q2A[i] = ISO ? df{0.f, 0.f} : df_add(df_mul(qxA, qxA), df_mul(qy, qy));
      echiA[i] = echi_g[(long long)(sg * SG + s) * N + q];
      arA[i].init();
      aiA[i].init();
      arB[i].init();
      aiB[i].init();
    }

    for (int b = b0; b < b1; ++b) {
      const int cur = (b - b0) & 1;
      wait_async_copies();
      __syncthreads();
      // per-BF operands are loaded before the next tile's copies are queued,
      // so waiting on them never waits on the prefetch
      const float kxh = gp.kx[b], kyh = gp.ky[b];
      const df kx{kxh, gp.kxl[b]}, ky{kyh, gp.kyl[b]};
      const float ak = gp.ak[b];
      float2 rowA = make_float2(1.f, 0.f);
      float2 colp[P];
      float2 colHalf = make_float2(1.f, 0.f);
      if (do_trotter) {
        rowA = rowph_g[(long long)b * N + q];
#pragma unroll
        for (int i = 0; i < P; ++i) colp[i] = colph_g[(long long)b * GCOLS + sg * SG + s0 + i * SSTEP];
        colHalf = colph_g[(long long)b * GCOLS + HALF];
      }
      float2 gHalf = make_float2(0.f, 0.f);
      if (!FROM_H && sg == 0) gHalf = src[((long long)b * GCOLS + HALF) * N + q];
      asm volatile("" ::: "memory");
      if (b + 1 < b1) issue(b + 1, cur ^ 1);
      if (b + 1 < b1) colterms(b + 1, cur ^ 1);
      if (FROM_H) fft_lds_il<N, SG, BLOCK, TSTRIDE, false>(abuf[cur], tileT, tw);
      const float2* tile = FROM_H ? tileT : abuf[cur];
      constexpr int TS = FROM_H ? TSTRIDE : N;
      if (!do_trotter) {
        if (Gt) {
          float2* gb = Gt + (long long)b * GCOLS * N;
#pragma unroll
          for (int i = 0; i < P; ++i) {
            const int s = s0 + i * SSTEP;
            const int j = sg * SG + s;
            const float2 z = tile[s * TS + q];
            if (j == 0) {
              const float2 zm = tile[s * TS + qm];
              gb[q] = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
              gb[(long long)HALF * N + q] = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
            } else {
              gb[(long long)j * N + q] = z;
            }
          }
        }
        continue;
      }
      // ISO row terms of d(Q -/+ k), -qx^2 +/- 2 qx kx, as float32 estimates
      // (the double-float terms are formed only for points near the soft edge)
      const df kx2{2.0f * kx.hi, 2.0f * kx.lo};
      float uq = 0.0f, vq = 0.0f;
      df cb{0.f, 0.f};
      if constexpr (ISO) {
        uq = row_estimate_minus(qxA, kx2, q2x);
        vq = row_estimate_plus(qxA, kx2, q2x);
      } else {
        cb = df_sub(rout2, df_add(df_mul(kx, kx), df_mul(ky, ky)));
      }
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int s = s0 + i * SSTEP;
        const int j = sg * SG + s;
        const float2 z = tile[s * TS + q];
        float2 GA, GB;
        if (j == 0) {
          if (FROM_H) {
            const float2 zm = tile[s * TS + qm];
            GA = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
            GB = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
          } else {
            GA = z;
            GB = gHalf;
          }
        } else {
          GA = z;
          GB = cconj(z);
        }
        const df qy{qyS[s], qySl[s]};
        float am, app;
        if constexpr (ISO) {
          iso_aperture_pair(uq, vq, qxA, kx2, q2x, colA_s[cur][s], colB_s[cur][s], ap, am, app);
        } else {
          aperture_pair<ISO>(qxA, qy, kx, ky, cb, q2A[i], ap, o, am, app);
        }
        const float2 eA = cmul(rowA, colp[i]);
        // Mirror points (j >= 1, q != N/2): output B sees the apertures
        // swapped, conj(e^{i theta}), the same e^{i chi} and conj(G), so its
        // Gamma is -conj(Gamma_A) with the same norm and its term is exactly
        // -conj(term_A) (every operation is sign-symmetric, Kahan included):
        // B is not accumulated and is written as -conj(A) at the end.
        const bool mirror = (j != 0) && (q != HALF);
        if (am != 0.0f || app != 0.0f) {
          const float wr = (am - app) * echiA[i].x;
          const float wi = -(am + app) * echiA[i].y;
          float gr = ak * (eA.x * wr - eA.y * wi);
          float gi = ak * (eA.x * wi + eA.y * wr);
          const float mag2 = gr * gr + gi * gi;
          const float sc = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
          gr *= sc;
          gi *= sc;
          arA[i].add(GA.x * gr + GA.y * gi);
          aiA[i].add(GA.y * gr - GA.x * gi);
        }
        // wave-uniform test first: most waves hold only mirror points and
        // branch over the B path
        const bool wave_b = __ballot(!mirror) != 0ull;
        if (wave_b) if (!mirror) {
          // j == 0: row q, column N/2; q == N/2: row N/2, column N-j
          const float2 echiB = (j == 0) ? ehalfc[q] : enyqr[s];
          float bm, bp;
          if constexpr (ISO) {
            // j == 0: row q, column N/2; q == N/2: row N/2 (this row), column -qy
            iso_aperture_pair(uq, vq, qxA, kx2, q2x, (j == 0) ? colA_s[cur][SG] : colB_s[cur][s],
                              (j == 0) ? colB_s[cur][SG] : colA_s[cur][s], ap, bm, bp);
          } else {
            const df qyB = (j == 0) ? df{qyS[SG], qySl[SG]} : df{-qy.hi, -qy.lo};
            const df qxB = (j == 0) ? qxA : qxB0;
            const df q2B = (j == 0) ? q2halfc[q] : q2A[i];
            aperture_pair<ISO>(qxB, qyB, kx, ky, cb, q2B, ap, o, bm, bp);
          }
          // rows q and -q coincide at q == N/2, so the row phasor is rowA
          const float2 eB = cmul(rowA, (j == 0) ? colHalf : cconj(colp[i]));
          if (bm != 0.0f || bp != 0.0f) {
            const float wr = (bm - bp) * echiB.x;
            const float wi = -(bm + bp) * echiB.y;
            float gr = ak * (eB.x * wr - eB.y * wi);
            float gi = ak * (eB.x * wi + eB.y * wr);
            const float mag2 = gr * gr + gi * gi;
            const float sc = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
            gr *= sc;
            gi *= sc;
            arB[i].add(GB.x * gr + GB.y * gi);
            aiB[i].add(GB.y * gr - GB.x * gi);
          }
        }
        if (Gt) {
          float2* gb = Gt + (long long)b * GCOLS * N;
          gb[(long long)j * N + q] = GA;
          if (j == 0) gb[(long long)HALF * N + q] = GB;
        }
        // one point at a time: keeps the scheduler from overlapping the
        // points' temporaries, which would spill the accumulators
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    if (do_trotter) {
      float2* sp = S_part + (long long)g * N * N;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int s = s0 + i * SSTEP;
        const int j = sg * SG + s;
        const int colB = (j == 0) ? HALF : N - j;
        const int rowB = (j == 0) ? q : qm;
        const float2 outA = make_float2(arA[i].s - arA[i].c, aiA[i].s - aiA[i].c);
        const bool mirror = (j != 0) && (q != HALF);
        sp[(long long)j * N + q] = outA;
        sp[(long long)colB * N + rowB] =
            mirror ? make_float2(-outA.x, outA.y) : make_float2(arB[i].s - arB[i].c, aiB[i].s - aiB[i].c);
      }
    }
  }
}

// ------------------------------------------------------------------------
// Wave-local fused column FFT + trotter of the reconstruction (row-transformed
// stack H, isotropic detector sampling).  A workgroup of NW waves takes one
// BF group and NW * CPW adjacent slot columns (8 waves x 2 columns: the 128
// bytes of every row of H that one cache line holds), and each wave owns CPW
// of those columns for every BF pixel of the group.  Lane l holds rows
// l + 64 m (m < R = N / 64): one 16- or 8-byte load per row brings the
// wave's columns.  The column FFT needs no workgroup barrier: an R-point DFT
// in registers (stride-64 elements) with twiddles w_N^(l k1), then 64-point
// Stockham radix-4 stages over the CPW R sequences through the wave's own
// LDS (every lane reads its operands before any lane writes), so
// X[k1 + R k2] = Z_k1[k2]; the last stage's outputs stay in the registers of
// the lane that computed them, lane l holding q = wave_qlane(l) + 64 r.
// Per point the trotter term is that of k_trotter_slots (same double-float
// apertures, phasors and guard), with Kahan FP32 sums over the group's BF
// pixels.  The aperture of a point far from the soft edge is decided from
// a float32 estimate of d (error < 2e-5 for |terms| < 32, inside the 1e-4
// margin of aperture_sum); only points near the edge evaluate d in
// double-float.  Outputs B of the mirror points are written as -conj(A);
// column N/2 (slot 0) is accumulated with A; the Nyquist-row outputs
// S(N/2, N - j) are summed after the BF loop, one BF pixel per lane, from
// the G_b(N/2, j) lane 0 holds.  The column terms of d for each (BF pixel,
// column) are computed once per workgroup into LDS, 64 BF pixels at a time;
// H and the phasors of the next BF pixel are loaded while this one is
// transformed.  Items (BF group, column group) start heavy-first on each
// XCD: the low slot columns hold more points inside the apertures.
// ------------------------------------------------------------------------
constexpr int kWaveColumnsDefault = 2;   // slot columns per wave of the default variant
constexpr int kWaveTrotterGroups = 64;
constexpr int kWaveChunk = 64;           // BF pixels per column-term table / Nyquist pass

__device__ __forceinline__ void wave_lds_fence() { asm volatile("" ::: "memory"); }

// Workgroup barrier for LDS only: this wave's LDS operations complete, then
// s_barrier.  Unlike __syncthreads it does not drain the wave's global loads,
// so prefetches of H stay in flight across it (their uses wait for them).
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }


// Column terms of d(Q -/+ k) for one (BF pixel, column): (Rout^2 - |k|^2) -
// qy^2 +/- 2 qy ky (isotropic sampling).
struct ColTerm {
  df A, B;
};

__device__ __forceinline__ ColTerm col_term(df rout2, df kx, df ky, df qy) {
  const df cbv = df_sub(rout2, df_add(df_mul(kx, kx), df_mul(ky, ky)));
  const df base = df_sub(cbv, df_mul(qy, qy));
  const df qk = df_mul(qy, df{2.0f * ky.hi, 2.0f * ky.lo});
  return ColTerm{df_add(base, qk), df_sub(base, qk)};
}

// After the wave-local column FFT, lane l holds frequencies
// q = wave_qlane(l) + 64 r (r = 0..3) of each of its columns: the outputs of the
// last Stockham stage stay in the registers of the lane that computed them.
// Per-row tables in LDS are stored in that order, entry 64 r + l for lane l.
__device__ __forceinline__ int wave_qlane(int lane) { return (lane >> 4) + 4 * (lane & 15); }

// Column c of one row's load: a float4 holds two adjacent slot columns.
__device__ __forceinline__ float2 column_of(float4 v, int c) {
  return c ? make_float2(v.z, v.w) : make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 column_of(float2 v, int) { return v; }

// Per-BF operands of one wave other than H (loaded one BF pixel ahead).
template <int R, int CPW>
struct WaveBf {
  float2 rowA[R];       // e^{i qx a_b} at the lane's frequencies
  float2 colp[CPW];     // e^{i qy b_b} at the wave's columns
  float2 colHalf;       // ... at column N/2
  float kx, kxl, ak;
};

template <int N, int R, int CPW>
__device__ __forceinline__ void wave_bf_load(WaveBf<R, CPW>& o, const float2* __restrict__ rowph_g,
                                             const float2* __restrict__ colph_g, const GeomPtrs& gp, int b,
                                             int lane, int j0) {
  constexpr int HALF = N / 2, GCOLS = HALF + 1;
#pragma unroll
  for (int m = 0; m < R; ++m) o.rowA[m] = rowph_g[(long long)b * N + wave_qlane(lane) + 64 * m];
#pragma unroll
  for (int c = 0; c < CPW; ++c) o.colp[c] = colph_g[(long long)b * GCOLS + j0 + c];
  o.colHalf = colph_g[(long long)b * GCOLS + HALF];
  o.kx = gp.kx[b];
  o.kxl = gp.kxl[b];
  o.ak = gp.ak[b];
}

// Normalised overlap weight and the term G * conj(g) of one point.
__device__ __forceinline__ float2 trotter_point(float2 G, float am, float app, float2 echi, float2 eth, float ak) {
  const float wr = (am - app) * echi.x;
  const float wi = -(am + app) * echi.y;
  float gr = ak * (eth.x * wr - eth.y * wi);
  float gi = ak * (eth.x * wi + eth.y * wr);
  const float mag2 = gr * gr + gi * gi;
  const float sc = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
  gr *= sc;
  gi *= sc;
  return make_float2(G.x * gr + G.y * gi, G.y * gr - G.x * gi);
}

// One wave's share of a work item: BF pixels [b0, b1) for the CPW slot
// columns j0, j0 + 1, ...; SLOT0 for the wave holding slot 0, whose first
// column packs the real columns 0 and N/2.  Every wave of the workgroup calls
// it with the same BF range (the barriers of the column-term tables match).
template <int N, int CPW, int NW, bool SLOT0>
__device__ __forceinline__ void wave_trotter_item(const float2* __restrict__ H, const GeomPtrs& gp,
                                                  const float2* __restrict__ rowph_g,
                                                  const float2* __restrict__ colph_g,
                                                  const float2* __restrict__ echi_g, const ApIso& ap, df rout2,
                                                  const float2* __restrict__ tw, float2* __restrict__ xb,
                                                  float2* __restrict__ nyqz,
                                                  ColTerm (*colT)[CPW * NW + 1],
                                                  const float2 (*echiS)[N], const df* __restrict__ qxS,
                                                  const df* __restrict__ q2xS, int cg, int w, int tid, int b0, int b1,
                                                  float2* __restrict__ sp) {
  constexpr int HALF = N / 2;
  constexpr int R = N / 64;
  constexpr int SP = 65;                 // padded sequence stride in the wave's LDS
  constexpr int CG = CPW * NW;
  constexpr int NBH = SLOT0 ? R : 1;     // column N/2 accumulators (slot 0 only)
  static_assert(CPW == 1 || CPW == 2, "one or two columns per wave");
  const int lane = tid & 63;
  const int j0 = cg * CG + CPW * w;
  Kahan aR[CPW][R], aI[CPW][R], hR[NBH], hI[NBH];
#pragma unroll
  for (int c = 0; c < CPW; ++c)
#pragma unroll
    for (int m = 0; m < R; ++m) {
      aR[c][m].init();
      aI[c][m].init();
    }
#pragma unroll
  for (int i = 0; i < NBH; ++i) {
    hR[i].init();
    hI[i].init();
  }
  // Nyquist-row sums (wave-uniform), value + running error term
  df nyR[CPW], nyI[CPW];
#pragma unroll
  for (int c = 0; c < CPW; ++c) {
    nyR[c] = df{0.f, 0.f};
    nyI[c] = df{0.f, 0.f};
  }
  // one load per row brings the wave's CPW columns (16 or 8 bytes)
  using HV = typename std::conditional<CPW == 2, float4, float2>::type;
  const HV* hrow = reinterpret_cast<const HV*>(H + (long long)lane * HALF + j0);
  constexpr long long ROWS64 = 64LL * HALF / CPW;   // HV between rows l and l + 64
  constexpr long long PER_B = (long long)N * HALF / CPW;
  HV h[R];
  // the per-BF operands other than H are loaded one BF pixel ahead with two
  // columns per wave; with one (a register budget of 168) at the top of the
  // BF pixel's iteration, where the column FFT covers their latency
  constexpr bool BF_AHEAD = CPW == 2;
  WaveBf<R, CPW> nx;
  if (b0 < b1) {
#pragma unroll
    for (int m = 0; m < R; ++m) h[m] = hrow[(long long)b0 * PER_B + m * ROWS64];
    if (BF_AHEAD) wave_bf_load<N, R, CPW>(nx, rowph_g, colph_g, gp, b0, lane, j0);
  }
  for (int c0 = b0; c0 < b1; c0 += kWaveChunk) {
    const int n = min(kWaveChunk, b1 - c0);
    // column terms of the chunk's BF pixels at the workgroup's CG columns
    // and column N/2, computed once for all waves
    lds_barrier();
    for (int e = tid; e < n * (CG + 1); e += 64 * NW) {
      const int bi = e / (CG + 1), ci = e - (e / (CG + 1)) * (CG + 1);
      const int b = c0 + bi;
      const int col = (ci < CG) ? cg * CG + ci : HALF;
      colT[bi][ci] = col_term(rout2, df{gp.kx[b], gp.kxl[b]}, df{gp.ky[b], gp.kyl[b]},
                              df{gp.qy[col], gp.qyl[col]});
    }
    lds_barrier();
    for (int b = c0; b < c0 + n; ++b) {
      WaveBf<R, CPW> cu;
      if constexpr (BF_AHEAD) {
        cu = nx;
        if (b + 1 < b1) wave_bf_load<N, R, CPW>(nx, rowph_g, colph_g, gp, b + 1, lane, j0);
      } else {
        wave_bf_load<N, R, CPW>(cu, rowph_g, colph_g, gp, b, lane, j0);
      }
      // column FFT: in-register radix-4 over m, twiddle, into the wave's LDS.
      // The exchange runs across the lanes of this wave only: a compiler
      // memory barrier at each phase boundary keeps every LDS access in
      // program order (the hardware keeps a wave's LDS operations in order).
      wave_lds_fence();
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        float2 x[R];
#pragma unroll
        for (int m = 0; m < R; ++m) x[m] = column_of(h[m], c);
        const float2 apc = cadd(x[0], x[2]), amc = csub(x[0], x[2]);
        const float2 bpd = cadd(x[1], x[3]), bmd = cmul_mi<false>(csub(x[1], x[3]));
        xb[(c * R + 0) * SP + lane] = cadd(apc, bpd);
        xb[(c * R + 1) * SP + lane] = cmul(cadd(amc, bmd), tw[lane & (N - 1)]);
        xb[(c * R + 2) * SP + lane] = cmul(csub(apc, bpd), tw[(2 * lane) & (N - 1)]);
        xb[(c * R + 3) * SP + lane] = cmul(csub(amc, bmd), tw[(3 * lane) & (N - 1)]);
      }
      // H of the next BF pixel: issued once this one's rows are consumed
      if (b + 1 < b1) {
#pragma unroll
        for (int m = 0; m < R; ++m) h[m] = hrow[(long long)(b + 1) * PER_B + m * ROWS64];
      }
      // 64-point Stockham radix-4 stages over the 2R sequences (two
      // butterflies per lane): all operands are read before any is written.
      // The last stage's butterfly p of lane l is column p at frequencies
      // q = wave_qlane(l) + 64 r; it stays in registers (slot 0's wave also
      // stores it, for the mirror rows N - q its column 0 needs).
      float2 zq[CPW][4];
#pragma unroll
      for (int ns = 1; ns < 64; ns *= 4) {
        wave_lds_fence();
        float2 v[CPW][4];
#pragma unroll
        for (int p = 0; p < CPW; ++p) {
          const int it = lane + 64 * p;
          const int sq = it >> 4, jj = it & 15, k = jj & (ns - 1);
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            float2 xv = xb[sq * SP + jj + 16 * r];
            if (ns > 1 && r > 0) xv = cmul(xv, tw[(4 * r * k * (16 / ns)) & (N - 1)]);
            v[p][r] = xv;
          }
        }
        wave_lds_fence();
#pragma unroll
        for (int p = 0; p < CPW; ++p) {
          const int it = lane + 64 * p;
          const int sq = it >> 4, jj = it & 15, k = jj & (ns - 1);
          const int base = sq * SP + (jj / ns) * ns * 4 + k;
          const float2 apc = cadd(v[p][0], v[p][2]), amc = csub(v[p][0], v[p][2]);
          const float2 bpd = cadd(v[p][1], v[p][3]), bmd = cmul_mi<false>(csub(v[p][1], v[p][3]));
          const float2 o[4] = {cadd(apc, bpd), cadd(amc, bmd), csub(apc, bpd), csub(amc, bmd)};
          if (ns < 16 || SLOT0) {
#pragma unroll
            for (int r = 0; r < 4; ++r) xb[base + r * ns] = o[r];
          }
          if (ns == 16) {
#pragma unroll
            for (int r = 0; r < 4; ++r) zq[p][r] = o[r];
          }
        }
      }
      wave_lds_fence();
      // G_b(N/2, j0 + c) for the deferred Nyquist pass (lane 0 holds q = N/2 at r = 2)
      static_assert(HALF == 128, "q = N/2 is lane 0's r = 2");
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < CPW; ++c) nyqz[CPW * (b - c0) + c] = zq[c][2];
      }
      ColTerm ct[CPW];
#pragma unroll
      for (int c = 0; c < CPW; ++c) ct[c] = colT[b - c0][CPW * w + c];
      const float ak = cu.ak;
      const df kx2{2.0f * cu.kx, 2.0f * cu.kxl};
      // LDS operands of the lane's points at q = wave_qlane(lane) + 64 m
      // (table entry 64 m + lane), read one row ahead so that no point waits
      // on its own reads
      struct RowOps {
        df qx, q2x;
        float2 e[CPW];
      };
      auto row_ops = [&](int m, RowOps& o) {
        const int i = lane + 64 * m;
        o.qx = qxS[i];
        o.q2x = q2xS[i];
#pragma unroll
        for (int c = 0; c < CPW; ++c) o.e[c] = echiS[CPW * w + c][i];
      };
      RowOps next_ops;
      row_ops(0, next_ops);
#pragma unroll
      for (int m = 0; m < R; ++m) {
        const int q = wave_qlane(lane) + 64 * m;
        const RowOps ops = next_ops;
        if (m + 1 < R) row_ops(m + 1, next_ops);
        const df qx = ops.qx;
        const df q2x = ops.q2x;
        const float uq = row_estimate_minus(qx, kx2, q2x);
        const float vq = row_estimate_plus(qx, kx2, q2x);
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
          const float2 z = zq[c][m];
          float2 GA;
          if (SLOT0 && c == 0) {
            const int qq = (N - q) & (N - 1);
            const float2 zm = xb[(qq & 3) * SP + (qq >> 2)];
            GA = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
          } else {
            GA = z;
          }
          float am, app;
          iso_aperture_pair(uq, vq, qx, kx2, q2x, ct[c].A, ct[c].B, ap, am, app);
          if (am != 0.0f || app != 0.0f) {
            const float2 t = trotter_point(GA, am, app, ops.e[c], cmul(cu.rowA[m], cu.colp[c]), ak);
            aR[c][m].add(t.x);
            aI[c][m].add(t.y);
          }
          if constexpr (SLOT0) {
            if (c == 0) {
              // output B of slot 0: column N/2 at row q
              const int qq = (N - q) & (N - 1);
              const float2 zm = xb[(qq & 3) * SP + (qq >> 2)];
              const float2 GB = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
              const ColTerm th = colT[b - c0][CG];
              float bm, bp;
              iso_aperture_pair(uq, vq, qx, kx2, q2x, th.A, th.B, ap, bm, bp);
              if (bm != 0.0f || bp != 0.0f) {
                const float2 t = trotter_point(GB, bm, bp, echiS[CG][lane + 64 * m], cmul(cu.rowA[m], cu.colHalf), ak);
                hR[m].add(t.x);
                hI[m].add(t.y);
              }
            }
          }
          // one point at a time: keeps the scheduler from overlapping the
          // points' temporaries
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
    // Nyquist-row terms of the chunk: lane i takes BF pixel c0 + i
    wave_lds_fence();
    {
      float2 t[CPW];
#pragma unroll
      for (int c = 0; c < CPW; ++c) t[c] = make_float2(0.f, 0.f);
      if (lane < n) {
        const int b = c0 + lane;
        const float ak = gp.ak[b];
        const df u = df_mul(qxS[HALF], df{2.0f * gp.kx[b], 2.0f * gp.kxl[b]});
        const df rowTA = df_sub(u, q2xS[HALF]);
        const df rowTB = df_neg(df_add(q2xS[HALF], u));
        const float2 rA = rowph_g[(long long)b * N + HALF];
#pragma unroll
        for (int c = 0; c < CPW; ++c) {
          const int j = j0 + c;
          if (j == 0) continue;          // slot 0: column N/2 holds this point
          const ColTerm ctn = colT[lane][CPW * w + c];
          const float bm = aperture_sum(rowTA, ctn.B, ap);
          const float bp = aperture_sum(rowTB, ctn.A, ap);
          if (bm != 0.0f || bp != 0.0f) {
            const float2 echiB = echi_g[(long long)((N - j) & (N - 1)) * N + HALF];
            const float2 eB = cmul(rA, cconj(colph_g[(long long)b * (HALF + 1) + j]));
            t[c] = trotter_point(cconj(nyqz[CPW * lane + c]), bm, bp, echiB, eB, ak);
          }
        }
      }
      // compensated xor tree over the lanes: two_sum's error term is exact,
      // so every lane holds the same (sum, error) pair
#pragma unroll
      for (int c = 0; c < CPW; ++c) {
        df re{t[c].x, 0.f}, im{t[c].y, 0.f};
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          const df ro{__shfl_xor(re.hi, off, 64), __shfl_xor(re.lo, off, 64)};
          const df io{__shfl_xor(im.hi, off, 64), __shfl_xor(im.lo, off, 64)};
          const df rs = two_sum(re.hi, ro.hi), is = two_sum(im.hi, io.hi);
          re = df{rs.hi, (re.lo + ro.lo) + rs.lo};
          im = df{is.hi, (im.lo + io.lo) + is.lo};
        }
        const df rs = two_sum(nyR[c].hi, re.hi), is = two_sum(nyI[c].hi, im.hi);
        nyR[c] = df{rs.hi, (nyR[c].lo + re.lo) + rs.lo};
        nyI[c] = df{is.hi, (nyI[c].lo + im.lo) + is.lo};
      }
    }
    wave_lds_fence();
  }
#pragma unroll
  for (int m = 0; m < R; ++m) {
    const int q = wave_qlane(lane) + 64 * m;
    const int qm = (N - q) & (N - 1);
#pragma unroll
    for (int c = 0; c < CPW; ++c) {
      const int j = j0 + c;
      const float2 outA = make_float2(aR[c][m].s - aR[c][m].c, aI[c][m].s - aI[c][m].c);
      sp[(long long)j * N + q] = outA;
      if (j == 0) {
        if constexpr (SLOT0) sp[(long long)HALF * N + q] = make_float2(hR[m].s - hR[m].c, hI[m].s - hI[m].c);
      } else if (q != HALF) {
        sp[(long long)(N - j) * N + qm] = make_float2(-outA.x, outA.y);
      } else {
        sp[(long long)(N - j) * N + HALF] = make_float2(nyR[c].hi + nyR[c].lo, nyI[c].hi + nyI[c].lo);
      }
    }
  }
}

template <int N, int CPW, int NW, int WAVES_PER_SIMD>
__global__ void __launch_bounds__(64 * NW, WAVES_PER_SIMD)
k_trotter_wave(const float2* __restrict__ H, const BfInfo* __restrict__ info, GeomPtrs gp,
               const float2* __restrict__ rowph_g, const float2* __restrict__ colph_g,
               const float2* __restrict__ echi_g, ApIso ap, df rout2, const float2* __restrict__ tw_g, int groups,
               float2* __restrict__ S_part, int part, int parts) {
  constexpr int HALF = N / 2;
  constexpr int R = N / 64;
  static_assert(R == 4, "the in-register pass is radix 4");
  constexpr int W = NW;
  constexpr int CG = CPW * W;            // slot columns per workgroup
  constexpr int NCG = HALF / CG;
  // LG workgroups share the 16 slot columns of one 128-byte line of H
  constexpr int LG = 16 / CG;
  constexpr int NL = NCG / LG;           // line groups (8)
  static_assert(CG <= 16 && 16 % CG == 0 && NL == 8, "workgroup columns must tile a line of H");
  constexpr int SP = 65;
  __shared__ float2 tw[N];
  __shared__ float2 xbuf[W][CPW * R * SP];
  __shared__ float2 nyqz[W][CPW * kWaveChunk];
  __shared__ ColTerm colT[kWaveChunk][CG + 1];
  __shared__ float2 echiS[CG + 1][N];    // e^{i chi} of the CG columns and of column N/2
  __shared__ df qxS[N], q2xS[N];
  const int tid = threadIdx.x;
  const int w = tid >> 6;
  // qxS, q2xS and echiS in lane order: entry 64 r + l is frequency
  // wave_qlane(l) + 64 r (entry N/2 is q = N/2 itself)
  for (int i = tid; i < N; i += 64 * W) {
    tw[i] = tw_g[i];
    const int q = wave_qlane(i & 63) + (i & ~63);
    const df qx{gp.qx[q], gp.qxl[q]};
    qxS[i] = qx;
    q2xS[i] = df_mul(qx, qx);
  }
  // part `part` of `parts`: BF pixels [lo, hi) into partial planes
  // [part * gpp, (part + 1) * gpp), gpp = groups / parts
  const int nbf_all = info->nbf;
  const int lo = (int)((long long)nbf_all * part / parts);
  const int hi = (int)((long long)nbf_all * (part + 1) / parts);
  const int gpp = groups / parts;
  const long long items = (long long)gpp * NCG;
  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    // Workgroups go round-robin over the 8 XCDs and, within one, start in
    // order as CUs free up.  XCD x takes the BF groups x, x + 8, ...; its
    // items run line by line of H in increasing slot order -- the low slot
    // columns hold more points inside the apertures, so the long items start
    // first and the short ones fill in at the end -- and the LG items that
    // share a line of H are consecutive (read once per L2).
    const int xcd = (int)(item & 7);
    const long long k8 = item >> 3;
    const int gx = gpp / 8;                // BF groups per XCD
    const int line = (int)(k8 / (gx * LG));
    const int rem = (int)(k8 % (gx * LG));
    const int gl = (rem / LG) * 8 + xcd;
    const int cg = line * LG + rem % LG;
    const int g = part * gpp + gl;
    int b0, b1;
    group_range(gl, gpp, hi - lo, b0, b1);
    b0 += lo;
    b1 += lo;
    __syncthreads();
    {
      // all of a thread's loads are issued before its LDS stores
      constexpr int FILL = ((CG + 1) * N + 64 * W - 1) / (64 * W);
      float2 ev[FILL];
#pragma unroll
      for (int k = 0; k < FILL; ++k) {
        const int i = tid + k * 64 * W;
        if (i < (CG + 1) * N) {
          const int c = i / N, e = i - (i / N) * N;
          const int col = (c < CG) ? cg * CG + c : HALF;
          ev[k] = echi_g[(long long)col * N + wave_qlane(e & 63) + (e & ~63)];
        }
      }
#pragma unroll
      for (int k = 0; k < FILL; ++k) {
        const int i = tid + k * 64 * W;
        if (i < (CG + 1) * N) echiS[i / N][i - (i / N) * N] = ev[k];
      }
    }
    // (the first column-term table's barrier orders these writes)
    float2* sp = S_part + (long long)g * N * N;
    if (cg == 0 && w == 0)
      wave_trotter_item<N, CPW, NW, true>(H, gp, rowph_g, colph_g, echi_g, ap, rout2, tw, xbuf[w], nyqz[w], colT, echiS, qxS,
                                 q2xS, cg, w, tid, b0, b1, sp);
    else
      wave_trotter_item<N, CPW, NW, false>(H, gp, rowph_g, colph_g, echi_g, ap, rout2, tw, xbuf[w], nyqz[w], colT, echiS, qxS,
                                  q2xS, cg, w, tid, b0, b1, sp);
  }
}

// ------------------------------------------------------------------------
// Trotter from a full (non-Hermitian) stack Gt[b][c][q], c < N: each full
// column is its own output (test seam for imported full-plane stacks).
// ------------------------------------------------------------------------
template <int N, int CG, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_trotter_full(const float2* __restrict__ Gt, const BfInfo* __restrict__ info, GeomPtrs gp,
               const float2* __restrict__ rowph_g, QuadChi m, Optics o, int groups, float2* __restrict__ S_part) {
  constexpr int NCG = N / CG;
  constexpr int P = CG * N / BLOCK;
  __shared__ float2 colph[CG];
  __shared__ float qyS[CG], qySl[CG];
  const int tid = threadIdx.x;
  const int nbf = info->nbf;
  const long long items = (long long)groups * NCG;
  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    const int g = (int)(item / NCG);
    const int cg = (int)(item - (long long)g * NCG);
    int b0, b1;
    group_range(g, groups, nbf, b0, b1);
    __syncthreads();
    if (tid < CG) {
      qyS[tid] = gp.qy[cg * CG + tid];
      qySl[tid] = gp.qyl[cg * CG + tid];
    }
    __syncthreads();
    float2 echi[P];
    Kahan ar[P], ai[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pidx = tid + i * BLOCK;
      const int row = pidx % N, cl = pidx / N;
      float sn, cs;
      sincosf(chi_quad(m, gp.qx[row], qyS[cl]), &sn, &cs);
      echi[i] = make_float2(cs, sn);
      ar[i].init();
      ai[i].init();
    }
    for (int b = b0; b < b1; ++b) {
      const Coord kc{gp.kx[b], gp.kxl[b], gp.ky[b], gp.kyl[b]};
      const float tb = 2.0f * (m.mxy * kc.x + m.myy * kc.y);
      __syncthreads();
      if (tid < CG) {
        float sn, cs;
        sincosf(qyS[tid] * tb, &sn, &cs);
        colph[tid] = make_float2(cs, sn);
      }
      __syncthreads();
      const float2* gb = Gt + (long long)b * N * N;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pidx = tid + i * BLOCK;
        const int row = pidx % N, cl = pidx / N;
        const int col = cg * CG + cl;
        const float2 G = gb[(long long)col * N + row];
        const Coord qc{gp.qx[row], gp.qxl[row], qyS[cl], qySl[cl]};
        const float2 eth = cmul(rowph_g[(long long)b * N + row], colph[cl]);
        const float2 t = trotter_term(G, qc, kc, gp.ak[b], eth, echi[i], o);
        ar[i].add(t.x);
        ai[i].add(t.y);
      }
    }
    float2* sp = S_part + (long long)g * N * N;
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int pidx = tid + i * BLOCK;
      const int row = pidx % N, cl = pidx / N;
      sp[(long long)(cg * CG + cl) * N + row] = make_float2(ar[i].s - ar[i].c, ai[i].s - ai[i].c);
    }
  }
}

// ------------------------------------------------------------------------
// Finish: reduce partial planes, inverse FFT, angle.
// ------------------------------------------------------------------------
// Sum of the groups' partial planes (Kahan, in group order), /N_bf, DC :=
// dc, then the inverse FFT along q of COUNT columns of column-major S[c][q]
template <int N, int COUNT>
__global__ void __launch_bounds__(256)
k_reduce_ifft_cols(const float2* __restrict__ S_part, const BfInfo* __restrict__ info, int groups,
                   const float2* __restrict__ tw_g, float2* __restrict__ S) {
  constexpr int STRIDE = N + 1;
  constexpr long long NN = (long long)N * N;
  __shared__ float2 buf[COUNT * STRIDE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  const int c0 = blockIdx.x * COUNT;
  const int nbf = info->nbf;
  const float inv = 1.0f / (float)(nbf > 0 ? nbf : 1);
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ci = idx / N, r = idx % N;
    const long long e = (long long)(c0 + ci) * N + r;
    Kahan kr, ki;
    kr.init();
    ki.init();
    int g = 0;
    for (; g + 16 <= groups; g += 16) {    // sixteen loads in flight, then the adds in group order
      float2 v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = S_part[(long long)(g + u) * NN + e];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        kr.add(v[u].x);
        ki.add(v[u].y);
      }
    }
    for (; g < groups; ++g) {
      const float2 v = S_part[(long long)g * NN + e];
      kr.add(v.x);
      ki.add(v.y);
    }
    buf[ci * STRIDE + r] = (e == 0) ? make_float2(info->dc_re, info->dc_im)
                                    : make_float2((kr.s - kr.c) * inv, (ki.s - ki.c) * inv);
  }
  __syncthreads();
  fft_lds<N, COUNT, 256, STRIDE, true>(buf, tw);
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ci = idx / N, r = idx % N;
    S[(long long)(c0 + ci) * N + r] = buf[ci * STRIDE + r];
  }
}

// inverse FFT along c for COUNT rows, scale 1/N^2, object + phase (row-major)
template <int N, int COUNT>
__global__ void __launch_bounds__(256)
k_ifft_rows_phase(const float2* __restrict__ S, const float2* __restrict__ tw_g, float2* __restrict__ obj,
                  float* __restrict__ phase) {
  constexpr int STRIDE = N + 1;
  __shared__ float2 buf[COUNT * STRIDE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  const int r0 = blockIdx.x * COUNT;
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int c = idx / COUNT, ri = idx % COUNT;
    buf[ri * STRIDE + c] = S[(long long)c * N + r0 + ri];
  }
  __syncthreads();
  fft_lds<N, COUNT, 256, STRIDE, true>(buf, tw);
  const float scale = 1.0f / ((float)N * (float)N);
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ri = idx / N, c = idx % N;
    const float2 v = cscale(buf[ri * STRIDE + c], scale);
    const long long o = (long long)(r0 + ri) * N + c;
    if (obj) obj[o] = v;
    if (phase) phase[o] = atan2f(v.y, v.x);
  }
}

// ------------------------------------------------------------------------
// Launch wrappers
// ------------------------------------------------------------------------
static unsigned grid_items(long long items, long long cap) {
  if (items < 1) items = 1;
  if (items > cap) items = cap;
  return (unsigned)items;
}

ApIso iso_aperture(const Optics& o, df& rout2_df) {
  // Rout = (semi + ang/2) / lambda; the isotropic form needs ang_y == ang_x
  const double lam = (double)o.lam.hi + (double)o.lam.lo;
  const double semi = (double)o.semi.hi + (double)o.semi.lo;
  const double ang = (double)o.angy.hi + (double)o.angy.lo;
  const double rout = (semi + 0.5 * ang) / lam;
  const double rin = (semi - 0.5 * ang) / lam;
  const double r2 = rout * rout;
  const float hi = (float)r2;
  rout2_df = df{hi, (float)(r2 - (double)hi)};
  ApIso a;
  a.rout2 = hi;
  a.rout = (float)rout;
  a.scale = (float)(lam / ang);
  // exact 0/1 outside the band, with a relative margin above FP32 noise
  a.d_in = (rin > 0.0) ? (float)((r2 - rin * rin) * (1.0 + 1e-5)) : INFINITY;
  a.d_out = (float)(-r2 * 1e-5);
  return a;
}

bool optics_isotropic(const Optics& o) { return o.angy.hi == o.angx.hi && o.angy.lo == o.angx.lo; }

hipError_t launch_phasors(int N, const BfInfo* info, int nbf_hint, const GeomPtrs& gp, QuadChi m, float2* rowph,
                          float2* colph, float2* echi, hipStream_t s,
                          const QuadChi* m_dev) {
  long long total = (long long)(nbf_hint > 0 ? nbf_hint : 1) * (N + N / 2 + 1);
  if (total < (long long)N * N) total = (long long)N * N;
  hipLaunchKernelGGL(k_phasors, dim3(grid_items((total + 255) / 256, 4096)), dim3(256), 0, s, info, gp, m, N, rowph,
                     colph, echi, m_dev);
  return hipGetLastError();
}

// Tile shapes (SG slot columns, BLOCK threads) of the fused column FFT +
// trotter on the row-transformed stack (the reconstruction path); variant 0
// is the default, the others are the tuner's candidates.  SG is even there:
// each lane's 16-byte staging copy moves two adjacent slots of a row of H.
// At N = 256 variants 0 and 1 are the wave-local kernel (the object path of
// a reconstruction: row-transformed stack, isotropic sampling) with
// kWaveColumnsDefault columns per wave and with the other count, and 2..4
// are the workgroup-tile shapes below; elsewhere the tile shapes only.
int trotter_variants(int N) { return N == 1024 ? 1 : (N == 256 ? 5 : 3); }

static int wave_columns(int variant) { return variant == 0 ? kWaveColumnsDefault : 3 - kWaveColumnsDefault; }

int trotter_groups_max(int N) { return N == 256 ? kWaveTrotterGroups : object_groups(N); }

bool trotter_wave_path(int N, bool from_h, const Optics& o, bool do_trotter, bool has_gt, int variant) {
  return N == 256 && from_h && optics_isotropic(o) && do_trotter && !has_gt && (variant == 0 || variant == 1);
}

hipError_t launch_trotter_slots(int N, bool from_h, const float2* src, const BfInfo* info, const GeomPtrs& gp,
                                const float2* rowph, const float2* colph, const float2* echi, const Optics& o,
                                const float2* tw, float2* S_part, float2* Gt, bool do_trotter, int variant,
                                hipStream_t s, int* groups_used, int part, int parts) {
  df rout2;
  const ApIso ap = iso_aperture(o, rout2);
  const int dt = do_trotter ? 1 : 0;
  const bool iso = optics_isotropic(o);
  if (trotter_wave_path(N, from_h, o, do_trotter, Gt != nullptr, variant)) {
    constexpr int groups = kWaveTrotterGroups;
    // (the item order gives each XCD groups / parts / 8 BF groups)
    if (parts < 1 || part < 0 || part >= parts || groups % (8 * parts)) return hipErrorInvalidValue;
    // one column per wave in 4-wave workgroups (3 per CU at <= 168 VGPRs),
    // or two per wave in 8-wave workgroups (one per CU)
    const int cpw = wave_columns(variant);
    const int nw = cpw == 1 ? 4 : 8;
    const unsigned grid = grid_items((long long)(groups / parts) * ((256 / 2) / (cpw * nw)), 65536);
    if (cpw == 1)
      hipLaunchKernelGGL((k_trotter_wave<256, 1, 4, 3>), dim3(grid), dim3(64 * 4), 0, s, src, info, gp, rowph, colph,
                         echi, ap, rout2, tw, groups, S_part, part, parts);
    else
      hipLaunchKernelGGL((k_trotter_wave<256, 2, 8, 2>), dim3(grid), dim3(64 * 8), 0, s, src, info, gp, rowph, colph,
                         echi, ap, rout2, tw, groups, S_part, part, parts);
    if (groups_used) *groups_used = groups;
    return hipGetLastError();
  }
  if (parts != 1) return hipErrorInvalidValue;
  // workgroup-tile shapes: at N = 256 their variants follow the wave-local ones
  if (N == 256 && variant > 0) variant = variant >= 2 ? variant - 2 : 0;
  // a resident stack at N = 256: 32 BF groups (2048 items, two rounds of four
  // workgroups per CU, so the heavy-first order can even out the CUs)
  const int groups = (!from_h && N == 256) ? 2 * object_groups(N) : object_groups(N);
  if (groups_used) *groups_used = groups;
#define QSSB_SLOTS_ONE(NN, SGV, BL, FH, IS)                                                                 \
  hipLaunchKernelGGL((k_trotter_slots<NN, SGV, BL, FH, IS>), dim3(grid), dim3(BL), 0, s, src, info, gp, rowph, \
                     colph, echi, ap, rout2, o, tw, groups, S_part, Gt, dt)
#define QSSB_SLOTS(NN, SGV, BL)                                                                \
  do {                                                                                        \
    const unsigned grid = grid_items((long long)groups * ((NN / 2) / SGV), 65536);            \
    if (from_h && iso) QSSB_SLOTS_ONE(NN, SGV, BL, true, true);                               \
    else if (from_h) QSSB_SLOTS_ONE(NN, SGV, BL, true, false);                                \
    else if (iso) QSSB_SLOTS_ONE(NN, SGV, BL, false, true);                                   \
    else QSSB_SLOTS_ONE(NN, SGV, BL, false, false);                                           \
  } while (0)
#define QSSB_SLOTS_TUNED(NN, SGV, BL)                                                          \
  do {                                                                                        \
    const unsigned grid = grid_items((long long)groups * ((NN / 2) / SGV), 65536);            \
    QSSB_SLOTS_ONE(NN, SGV, BL, true, true);                                                  \
  } while (0)
  const bool tuned = from_h && iso && variant > 0;
  switch (N) {
    case 128:
      if (tuned && variant == 1) QSSB_SLOTS_TUNED(128, 2, 256);
      else if (tuned && variant == 2) QSSB_SLOTS_TUNED(128, 8, 256);
      else QSSB_SLOTS(128, 4, 256);
      break;
    case 256:
      if (tuned && variant == 1) QSSB_SLOTS_TUNED(256, 4, 256);
      else if (tuned && variant == 2) QSSB_SLOTS_TUNED(256, 2, 512);
      else QSSB_SLOTS(256, 2, 256);
      break;
    case 512:
      if (tuned && variant == 1) QSSB_SLOTS_TUNED(512, 2, 1024);
      else if (tuned && variant == 2) QSSB_SLOTS_TUNED(512, 4, 512);
      else QSSB_SLOTS(512, 2, 512);
      break;
    case 1024:
      if (from_h) {
        // one slot per tile would split the 16-byte staging copies across rows
        const unsigned grid = grid_items((long long)groups * ((1024 / 2) / 2), 65536);
        if (iso) QSSB_SLOTS_ONE(1024, 2, 1024, true, true);
        else QSSB_SLOTS_ONE(1024, 2, 1024, true, false);
      } else {
        const unsigned grid = grid_items((long long)groups * ((1024 / 2) / 1), 65536);
        if (iso) QSSB_SLOTS_ONE(1024, 1, 1024, false, true);
        else QSSB_SLOTS_ONE(1024, 1, 1024, false, false);
      }
      break;
    default:
      return hipErrorInvalidValue;
  }
#undef QSSB_SLOTS_TUNED
#undef QSSB_SLOTS
#undef QSSB_SLOTS_ONE
  return hipGetLastError();
}

hipError_t launch_trotter_full(int N, const float2* Gt, const BfInfo* info, const GeomPtrs& gp, const float2* rowph,
                               QuadChi m, const Optics& o, float2* S_part, hipStream_t s) {
  const int groups = object_groups(N);
  switch (N) {
    case 128:
      hipLaunchKernelGGL((k_trotter_full<128, 16, 256>), dim3(grid_items(groups * 8LL, 65536)), dim3(256), 0, s, Gt,
                         info, gp, rowph, m, o, groups, S_part);
      break;
    case 256:
      hipLaunchKernelGGL((k_trotter_full<256, 8, 256>), dim3(grid_items(groups * 32LL, 65536)), dim3(256), 0, s, Gt,
                         info, gp, rowph, m, o, groups, S_part);
      break;
    case 512:
      hipLaunchKernelGGL((k_trotter_full<512, 4, 256>), dim3(grid_items(groups * 128LL, 65536)), dim3(256), 0, s,
                         Gt, info, gp, rowph, m, o, groups, S_part);
      break;
    case 1024:
      hipLaunchKernelGGL((k_trotter_full<1024, 1, 256>), dim3(grid_items(groups * 1024LL, 65536)), dim3(256), 0, s,
                         Gt, info, gp, rowph, m, o, groups, S_part);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_finish_object(int N, const float2* S_part, int groups, const BfInfo* info, const float2* tw,
                                float2* S, float2* obj, float* phase, hipStream_t s) {
  const int G = groups;
  switch (N) {
    case 128:
      hipLaunchKernelGGL((k_reduce_ifft_cols<128, 4>), dim3(128 / 4), dim3(256), 0, s, S_part, info, G, tw, S);
      hipLaunchKernelGGL((k_ifft_rows_phase<128, 8>), dim3(128 / 8), dim3(256), 0, s, S, tw, obj, phase);
      break;
    case 256:
      hipLaunchKernelGGL((k_reduce_ifft_cols<256, 1>), dim3(256), dim3(256), 0, s, S_part, info, G, tw, S);
      hipLaunchKernelGGL((k_ifft_rows_phase<256, 4>), dim3(256 / 4), dim3(256), 0, s, S, tw, obj, phase);
      break;
    case 512:
      hipLaunchKernelGGL((k_reduce_ifft_cols<512, 1>), dim3(512), dim3(256), 0, s, S_part, info, G, tw, S);
      hipLaunchKernelGGL((k_ifft_rows_phase<512, 2>), dim3(512 / 2), dim3(256), 0, s, S, tw, obj, phase);
      break;
    case 1024:
      hipLaunchKernelGGL((k_reduce_ifft_cols<1024, 1>), dim3(1024), dim3(256), 0, s, S_part, info, G, tw, S);
      hipLaunchKernelGGL((k_ifft_rows_phase<1024, 1>), dim3(1024), dim3(256), 0, s, S, tw, obj, phase);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qssb
