// Object-path kernels of the MI355X SSB engine: the scan-column FFT fused
// with the trotter reduction over bright-field pixels, the same reduction
// from a resident Fourier stack, and the final inverse FFT to the phase.
//
// Reference behaviour (paths relative to bobleesj/quantem.gpu):
//   Fourier-sum trotter   src/quantem/gpu/ssb/compute/cuda/kernels/fft256.py:967-1030
//   overlap / guard       src/quantem/gpu/ssb/compute/cuda/kernels/common.py:115-236
//   object + IFFT         src/quantem/gpu/ssb/compute/cuda/engine.py:1047-1107
//   phase = angle(object) src/quantem/gpu/ssb/results.py:117-122
#include "ssb_internal.h"

namespace qssb {

// BF pixels per partial plane of the trotter reduction.
constexpr int BG = 16;

int group_count(int N, int nbf) {
  (void)N;
  return (nbf + BG - 1) / BG;
}

__device__ __forceinline__ float chi_quad(const QuadChi& m, float x, float y) {
  return m.mxx * x * x + 2.0f * m.mxy * x * y + m.myy * y * y;
}

// ------------------------------------------------------------------------
// Scan-column FFT of H + trotter.  Work item = (BF group g, slot group sg).
// For each BF pixel of the group the SG slot columns of H[b] (N rows each)
// are staged in LDS, transformed along R_r, and every resulting G_b(q, j)
// contributes to S(q, j) and, through Hermitian symmetry, to
// S(-q, N-j) = sum_b conj(G_b(q, j)) conj(Gamma^_b(-q, N-j)).  Slot 0 packs
// the two real columns 0 and N/2 and is separated after the FFT.
// BLOCK is a multiple of N, so each thread keeps one scan frequency row.
// ------------------------------------------------------------------------
template <int N, int SG, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_colfft_trotter(const float2* __restrict__ H, const BfInfo* __restrict__ info, GeomPtrs gp, QuadChi m,
                 Optics o, const float2* __restrict__ tw_g, float2* __restrict__ S_part,
                 float2* __restrict__ Gt, int do_trotter) {
  constexpr int STRIDE = N + 1;
  constexpr int HALF = N / 2;
  constexpr int NSG = HALF / SG;
  constexpr int P = SG * N / BLOCK;       // points per thread
  constexpr int SSTEP = BLOCK / N;        // slot step between a thread's points
  constexpr int GCOLS = HALF + 1;
  __shared__ float2 tile[SG * STRIDE];
  __shared__ float2 tw[N];
  __shared__ float2 colph[2 * SG];
  __shared__ float qyS[2 * SG];
  __shared__ float qySl[2 * SG];

  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BLOCK) tw[i] = tw_g[i];
  const int nbf = info->nbf;
  const float2 dc = make_float2(info->dc_re, info->dc_im);
  (void)dc;
  const int ngroups = (nbf + BG - 1) / BG;
  const long long items = (long long)ngroups * NSG;
  const int q = tid % N;
  const int s0 = tid / N;
  const int qm = (N - q) & (N - 1);
  const float qxA = gp.qx[q], qxAl = gp.qxl[q];
  const float qxBm = gp.qx[qm], qxBml = gp.qxl[qm];

  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    const int g = (int)(item / NSG);
    const int sg = (int)(item - (long long)g * NSG);
    const int b0 = g * BG;
    const int b1 = min(nbf, b0 + BG);
    __syncthreads();
    if (tid < 2 * SG) {
      const int s = tid % SG;
      const int j = sg * SG + s;
      int col;
      if (tid < SG) col = j;
      else col = (j == 0) ? HALF : N - j;
      qyS[tid] = gp.qy[col];
      qySl[tid] = gp.qyl[col];
    }
    __syncthreads();
    // per-output constants: e^{i chi(Q)} for outputs A and B of each point
    float2 echiA[P], echiB[P];
    Kahan arA[P], aiA[P], arB[P], aiB[P];
#pragma unroll
    for (int i = 0; i < P; ++i) {
      const int s = s0 + i * SSTEP;
      const int j = sg * SG + s;
      const float qxB = (j == 0) ? qxA : qxBm;
      float sn, cs;
      sincosf(chi_quad(m, qxA, qyS[s]), &sn, &cs);
      echiA[i] = make_float2(cs, sn);
      sincosf(chi_quad(m, qxB, qyS[SG + s]), &sn, &cs);
      echiB[i] = make_float2(cs, sn);
      arA[i].init();
      aiA[i].init();
      arB[i].init();
      aiB[i].init();
    }
    for (int b = b0; b < b1; ++b) {
      __syncthreads();
      const float2* hb = H + (long long)b * N * HALF + sg * SG;
      for (int idx = tid; idx < N * SG; idx += BLOCK) {
        const int s = idx % SG;
        const int r = idx / SG;
        tile[s * STRIDE + r] = hb[(long long)r * HALF + s];
      }
      const Coord kc{gp.kx[b], gp.kxl[b], gp.ky[b], gp.kyl[b]};
      const float kx = kc.x, ky = kc.y, ak = gp.ak[b];
      const float ta = 2.0f * (m.mxx * kx + m.mxy * ky);
      const float tb = 2.0f * (m.mxy * kx + m.myy * ky);
      if (do_trotter && tid < 2 * SG) {
        float sn, cs;
        sincosf(qyS[tid] * tb, &sn, &cs);
        colph[tid] = make_float2(cs, sn);
      }
      __syncthreads();
      fft_lds<N, SG, BLOCK, STRIDE, false>(tile, tw);
      float2 rowA = make_float2(1.f, 0.f), rowBc = rowA;
      if (do_trotter) {
        float sn, cs;
        sincosf(qxA * ta, &sn, &cs);
        rowA = make_float2(cs, sn);
        rowBc = (q == HALF || q == 0) ? rowA : cconj(rowA);
      }
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int s = s0 + i * SSTEP;
        const int j = sg * SG + s;
        float2 GA, GB;
        const float2 z = tile[s * STRIDE + q];
        if (j == 0) {
          const float2 zm = tile[s * STRIDE + qm];
          GA = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
          GB = make_float2(0.5f * (z.y + zm.y), -0.5f * (z.x - zm.x));
        } else {
          GA = z;
          GB = cconj(z);
        }
        if (Gt) {
          float2* gb = Gt + (long long)b * GCOLS * N;
          gb[(long long)j * N + q] = GA;
          if (j == 0) gb[(long long)HALF * N + q] = GB;
        }
        if (do_trotter) {
          const float2 rowB = (j == 0) ? rowA : rowBc;
          const Coord qa{qxA, qxAl, qyS[s], qySl[s]};
          const Coord qb{(j == 0) ? qxA : qxBm, (j == 0) ? qxAl : qxBml, qyS[SG + s], qySl[SG + s]};
          const float2 tA = trotter_term(GA, qa, kc, ak, cmul(rowA, colph[s]), echiA[i], o);
          const float2 tB = trotter_term(GB, qb, kc, ak, cmul(rowB, colph[SG + s]), echiB[i], o);
          arA[i].add(tA.x);
          aiA[i].add(tA.y);
          arB[i].add(tB.x);
          aiB[i].add(tB.y);
        }
      }
    }
    if (do_trotter) {
      float2* sp = S_part + (long long)g * N * N;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int s = s0 + i * SSTEP;
        const int j = sg * SG + s;
        const int colA = j;
        const int colB = (j == 0) ? HALF : N - j;
        const int rowB = (j == 0) ? q : qm;
        sp[(long long)colA * N + q] = make_float2(arA[i].s - arA[i].c, aiA[i].s - aiA[i].c);
        sp[(long long)colB * N + rowB] = make_float2(arB[i].s - arB[i].c, aiB[i].s - aiB[i].c);
      }
    }
  }
}

// ------------------------------------------------------------------------
// Trotter reduction from a resident stack Gt[b][c][q] (gcols columns).
// Work item = (BF group g, column group of CG full-plane columns).
// ------------------------------------------------------------------------
template <int N, int CG, int BLOCK>
__global__ void __launch_bounds__(BLOCK)
k_trotter_from_g(const float2* __restrict__ Gt, int gcols, const BfInfo* __restrict__ info, GeomPtrs gp,
                 QuadChi m, Optics o, float2* __restrict__ S_part) {
  constexpr int NCG = N / CG;
  constexpr int P = CG * N / BLOCK;
  __shared__ float2 rowph[N];
  __shared__ float2 colph[CG];
  __shared__ float qxs[N];
  __shared__ float qxsl[N];
  __shared__ float qyS[CG];
  __shared__ float qySl[CG];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += BLOCK) {
    qxs[i] = gp.qx[i];
    qxsl[i] = gp.qxl[i];
  }
  const int nbf = info->nbf;
  const int ngroups = (nbf + BG - 1) / BG;
  const long long items = (long long)ngroups * NCG;
  for (long long item = blockIdx.x; item < items; item += gridDim.x) {
    const int g = (int)(item / NCG);
    const int cg = (int)(item - (long long)g * NCG);
    const int b0 = g * BG;
    const int b1 = min(nbf, b0 + BG);
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
      sincosf(chi_quad(m, qxs[row], qyS[cl]), &sn, &cs);
      echi[i] = make_float2(cs, sn);
      ar[i].init();
      ai[i].init();
    }
    for (int b = b0; b < b1; ++b) {
      const Coord kc{gp.kx[b], gp.kxl[b], gp.ky[b], gp.kyl[b]};
      const float kx = kc.x, ky = kc.y, ak = gp.ak[b];
      const float ta = 2.0f * (m.mxx * kx + m.mxy * ky);
      const float tb = 2.0f * (m.mxy * kx + m.myy * ky);
      __syncthreads();
      for (int r = tid; r < N; r += BLOCK) {
        float sn, cs;
        sincosf(qxs[r] * ta, &sn, &cs);
        rowph[r] = make_float2(cs, sn);
      }
      if (tid < CG) {
        float sn, cs;
        sincosf(qyS[tid] * tb, &sn, &cs);
        colph[tid] = make_float2(cs, sn);
      }
      __syncthreads();
      const float2* gb = Gt + (long long)b * gcols * N;
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int pidx = tid + i * BLOCK;
        const int row = pidx % N, cl = pidx / N;
        const int col = cg * CG + cl;
        float2 G;
        if (col < gcols) {
          G = gb[(long long)col * N + row];
        } else {
          G = cconj(gb[(long long)(N - col) * N + ((N - row) & (N - 1))]);
        }
        const Coord qc{qxs[row], qxsl[row], qyS[cl], qySl[cl]};
        const float2 t = trotter_term(G, qc, kc, ak, cmul(rowph[row], colph[cl]), echi[i], o);
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
__global__ void __launch_bounds__(256)
k_reduce_partials(const float2* __restrict__ S_part, const BfInfo* __restrict__ info, int NN,
                  float2* __restrict__ S) {
  const int idx = blockIdx.x * 256 + threadIdx.x;
  if (idx >= NN) return;
  const int nbf = info->nbf;
  const int ngroups = (nbf + BG - 1) / BG;
  Kahan kr, ki;
  kr.init();
  ki.init();
  for (int g = 0; g < ngroups; ++g) {
    const float2 v = S_part[(long long)g * NN + idx];
    kr.add(v.x);
    ki.add(v.y);
  }
  float2 out;
  if (idx == 0) {
    out = make_float2(info->dc_re, info->dc_im);
  } else {
    const float inv = 1.0f / (float)(nbf > 0 ? nbf : 1);
    out = make_float2((kr.s - kr.c) * inv, (ki.s - ki.c) * inv);
  }
  S[idx] = out;
}

// inverse FFT along q of column-major S[c][q] (in place), COUNT columns per block
template <int N, int COUNT>
__global__ void __launch_bounds__(256)
k_ifft_cols(float2* __restrict__ S, const float2* __restrict__ tw_g) {
  constexpr int STRIDE = N + 1;
  __shared__ float2 buf[COUNT * STRIDE];
  __shared__ float2 tw[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += 256) tw[i] = tw_g[i];
  const int c0 = blockIdx.x * COUNT;
  for (int idx = tid; idx < COUNT * N; idx += 256) {
    const int ci = idx / N, r = idx % N;
    buf[ci * STRIDE + r] = S[(long long)(c0 + ci) * N + r];
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

int colfft_group_size(int N) { return (N == 512) ? 8 : 16; }

hipError_t launch_colfft_trotter(int N, const float2* H, const BfInfo* info, int nbf_hint, const GeomPtrs& gp,
                                 QuadChi m, const Optics& o, const float2* tw, float2* S_part, int max_groups,
                                 float2* Gt, bool do_trotter, hipStream_t s) {
  (void)max_groups;
  const long long ng = (nbf_hint + BG - 1) / BG;
  const int dt = do_trotter ? 1 : 0;
  switch (N) {
    case 128:
      hipLaunchKernelGGL((k_colfft_trotter<128, 16, 512>), dim3(grid_items(ng * 4, 8192)), dim3(512), 0, s, H,
                         info, gp, m, o, tw, S_part, Gt, dt);
      break;
    case 256:
      hipLaunchKernelGGL((k_colfft_trotter<256, 16, 512>), dim3(grid_items(ng * 8, 8192)), dim3(512), 0, s, H,
                         info, gp, m, o, tw, S_part, Gt, dt);
      break;
    case 512:
      hipLaunchKernelGGL((k_colfft_trotter<512, 8, 512>), dim3(grid_items(ng * 32, 8192)), dim3(512), 0, s, H,
                         info, gp, m, o, tw, S_part, Gt, dt);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_trotter_from_g(int N, int gcols, const float2* Gt, const BfInfo* info, int nbf_hint,
                                 const GeomPtrs& gp, QuadChi m, const Optics& o, float2* S_part, int max_groups,
                                 hipStream_t s) {
  (void)max_groups;
  const long long ng = (nbf_hint + BG - 1) / BG;
  switch (N) {
    case 128:
      hipLaunchKernelGGL((k_trotter_from_g<128, 16, 256>), dim3(grid_items(ng * 8, 8192)), dim3(256), 0, s, Gt,
                         gcols, info, gp, m, o, S_part);
      break;
    case 256:
      hipLaunchKernelGGL((k_trotter_from_g<256, 8, 256>), dim3(grid_items(ng * 32, 8192)), dim3(256), 0, s, Gt,
                         gcols, info, gp, m, o, S_part);
      break;
    case 512:
      hipLaunchKernelGGL((k_trotter_from_g<512, 4, 256>), dim3(grid_items(ng * 128, 8192)), dim3(256), 0, s, Gt,
                         gcols, info, gp, m, o, S_part);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_finish_object(int N, const float2* S_part, const BfInfo* info, int nbf_hint, int max_groups,
                                const float2* tw, float2* S, float2* obj, float* phase, hipStream_t s) {
  (void)nbf_hint;
  (void)max_groups;
  const int NN = N * N;
  hipLaunchKernelGGL(k_reduce_partials, dim3((NN + 255) / 256), dim3(256), 0, s, S_part, info, NN, S);
  switch (N) {
    case 128:
      hipLaunchKernelGGL((k_ifft_cols<128, 8>), dim3(128 / 8), dim3(256), 0, s, S, tw);
      hipLaunchKernelGGL((k_ifft_rows_phase<128, 8>), dim3(128 / 8), dim3(256), 0, s, S, tw, obj, phase);
      break;
    case 256:
      hipLaunchKernelGGL((k_ifft_cols<256, 4>), dim3(256 / 4), dim3(256), 0, s, S, tw);
      hipLaunchKernelGGL((k_ifft_rows_phase<256, 4>), dim3(256 / 4), dim3(256), 0, s, S, tw, obj, phase);
      break;
    case 512:
      hipLaunchKernelGGL((k_ifft_cols<512, 2>), dim3(512 / 2), dim3(256), 0, s, S, tw);
      hipLaunchKernelGGL((k_ifft_rows_phase<512, 2>), dim3(512 / 2), dim3(256), 0, s, S, tw, obj, phase);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace qssb
