// Device-side building blocks of the MI355X SSB engine (gfx950, wave64).
//
// * complex helpers on float2;
// * an LDS Stockham FFT (radix-4 stages plus one radix-2 stage when log2 N is
//   odd) that a whole workgroup runs over several independent sequences;
// * the soft probe aperture of the reference (cuda/kernels/common.py:115-171);
// * the SSB overlap term in factorised form (see trotter_term below).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qssb {

// ---------------------------------------------------------------- complex
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by -i (forward) or +i (inverse)
template <bool INV>
__device__ __forceinline__ float2 cmul_mi(float2 a) {
  return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// Kahan-compensated FP32 accumulator (one real component).
struct Kahan {
  float s, c;
  __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
  __device__ __forceinline__ void add(float x) {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

// ---------------------------------------------------------------- FFT
// Stockham autosort FFT executed by all BLOCK threads of a workgroup over
// COUNT sequences of length N held in LDS at buf[s * STRIDE + i].
// `tw` (LDS or global) holds exp(-2*pi*i*m/N), m = 0..N-1, computed on the
// host in double precision.  Each stage reads every operand into registers,
// synchronises, then writes its outputs, so a single buffer suffices.
// Unnormalised in both directions.
template <int N, int R, int NS, int COUNT, int BLOCK, int STRIDE, bool INV>
__device__ __forceinline__ void fft_stage(float2* __restrict__ buf, const float2* __restrict__ tw) {
  constexpr int ITEMS = COUNT * (N / R);
  constexpr int PER = (ITEMS + BLOCK - 1) / BLOCK;
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int it = tid + p * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int s = it / (N / R);
      const int j = it - s * (N / R);
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = buf[s * STRIDE + j + r * (N / R)];
        if (NS > 1 && r > 0) {
          float2 w = tw[(r * k * (N / (NS * R))) & (N - 1)];
          if (INV) w.y = -w.y;
          x = cmul(x, w);
        }
        v[p][r] = x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int it = tid + p * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int s = it / (N / R);
      const int j = it - s * (N / R);
      const int k = j % NS;
      const int base = s * STRIDE + (j / NS) * NS * R + k;
      if constexpr (R == 4) {
        float2 a = v[p][0], b = v[p][1], c = v[p][2], d = v[p][3];
        float2 apc = cadd(a, c), amc = csub(a, c);
        float2 bpd = cadd(b, d), bmd = cmul_mi<INV>(csub(b, d));
        buf[base] = cadd(apc, bpd);
        buf[base + NS] = cadd(amc, bmd);
        buf[base + 2 * NS] = csub(apc, bpd);
        buf[base + 3 * NS] = csub(amc, bmd);
      } else {
        float2 a = v[p][0], b = v[p][1];
        buf[base] = cadd(a, b);
        buf[base + NS] = csub(a, b);
      }
    }
  }
  __syncthreads();
}

template <int N, int NS, int COUNT, int BLOCK, int STRIDE, bool INV>
__device__ __forceinline__ void fft_stages(float2* __restrict__ buf, const float2* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int R = (N / NS >= 4) ? 4 : 2;
    fft_stage<N, R, NS, COUNT, BLOCK, STRIDE, INV>(buf, tw);
    fft_stages<N, NS * R, COUNT, BLOCK, STRIDE, INV>(buf, tw);
  }
}

// Full FFT of COUNT sequences.  Caller must __syncthreads() before (data
// written) -- the routine ends with a barrier.
template <int N, int COUNT, int BLOCK, int STRIDE, bool INV>
__device__ __forceinline__ void fft_lds(float2* __restrict__ buf, const float2* __restrict__ tw) {
  fft_stages<N, 1, COUNT, BLOCK, STRIDE, INV>(buf, tw);
}

// Same Stockham FFT for COUNT sequences stored interleaved in blocks of SB:
// element (s, i) at buf[(s / SB) * BSTR + i * SB + s % SB] -- SB adjacent
// columns of a row-major array copied row by row, COUNT / SB such tiles
// BSTR apart.  Work items run s fastest, so a wave touches contiguous LDS.
// The last stage writes to `out` in the padded per-sequence layout
// out[s * OSTRIDE + i] (a separate buffer).
template <int N, int R, int NS, int COUNT, int BLOCK, bool INV, int OSTRIDE, int SB, int BSTR>
__device__ __forceinline__ void fft_stage_il(float2* __restrict__ buf, float2* __restrict__ out,
                                             const float2* __restrict__ tw) {
  static_assert(COUNT % SB == 0, "sequences come in whole tiles");
  auto at = [](int i, int s) { return (s / SB) * BSTR + i * SB + (s % SB); };
  constexpr int ITEMS = COUNT * (N / R);
  constexpr int PER = (ITEMS + BLOCK - 1) / BLOCK;
  constexpr bool LAST = (NS * R == N);
  float2 v[PER][R];
  const int tid = threadIdx.x;
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int it = tid + p * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int s = it % COUNT;
      const int j = it / COUNT;
      const int k = j % NS;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 x = buf[at(j + r * (N / R), s)];
        if (NS > 1 && r > 0) {
          float2 w = tw[(r * k * (N / (NS * R))) & (N - 1)];
          if (INV) w.y = -w.y;
          x = cmul(x, w);
        }
        v[p][r] = x;
      }
    }
  }
  if (!LAST) __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int it = tid + p * BLOCK;
    if (ITEMS % BLOCK == 0 || it < ITEMS) {
      const int s = it % COUNT;
      const int j = it / COUNT;
      const int k = j % NS;
      const int base = (j / NS) * NS * R + k;
      float2 y[R];
      if constexpr (R == 4) {
        float2 a = v[p][0], b = v[p][1], c = v[p][2], d = v[p][3];
        float2 apc = cadd(a, c), amc = csub(a, c);
        float2 bpd = cadd(b, d), bmd = cmul_mi<INV>(csub(b, d));
        y[0] = cadd(apc, bpd);
        y[1] = cadd(amc, bmd);
        y[2] = csub(apc, bpd);
        y[3] = csub(amc, bmd);
      } else {
        y[0] = cadd(v[p][0], v[p][1]);
        y[1] = csub(v[p][0], v[p][1]);
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if (LAST) out[s * OSTRIDE + base + r * NS] = y[r];
        else buf[at(base + r * NS, s)] = y[r];
      }
    }
  }
  __syncthreads();
}

template <int N, int NS, int COUNT, int BLOCK, bool INV, int OSTRIDE, int SB, int BSTR>
__device__ __forceinline__ void fft_stages_il(float2* __restrict__ buf, float2* __restrict__ out,
                                              const float2* __restrict__ tw) {
  if constexpr (NS < N) {
    constexpr int R = (N / NS >= 4) ? 4 : 2;
    fft_stage_il<N, R, NS, COUNT, BLOCK, INV, OSTRIDE, SB, BSTR>(buf, out, tw);
    fft_stages_il<N, NS * R, COUNT, BLOCK, INV, OSTRIDE, SB, BSTR>(buf, out, tw);
  }
}

// FFT of interleaved sequences in `buf` (destroyed) with the result in `out`
// (padded per-sequence layout).  Caller synchronises before; ends with a barrier.
template <int N, int COUNT, int BLOCK, int OSTRIDE, bool INV, int SB = COUNT, int BSTR = 0>
__device__ __forceinline__ void fft_lds_il(float2* __restrict__ buf, float2* __restrict__ out,
                                           const float2* __restrict__ tw) {
  fft_stages_il<N, 1, COUNT, BLOCK, INV, OSTRIDE, SB, BSTR>(buf, out, tw);
}

// Asynchronous global -> LDS copy of 16 bytes per lane (gfx950
// global_load_lds_dwordx4): lane l of the wave writes lds_wave_base + 16*l.
// `lds_wave_base` must be wave-uniform.  Completion: wait_async_copies().
__device__ __forceinline__ void copy_async16(const void* gptr, void* lds_wave_base) {
  __builtin_amdgcn_global_load_lds(gptr, (__attribute__((address_space(3))) void*)lds_wave_base, 16, 0, 0);
}
__device__ __forceinline__ void wait_async_copies() { __builtin_amdgcn_s_waitcnt(0); }

// ---------------------------------------------------------------- double-float
// Unevaluated sums hi + lo of two floats (~44-bit significand), built from
// error-free transformations (Dekker / Knuth two-sum, FMA two-product).
// Used only where FP32 cancellation decides a discontinuous outcome.
// Every routine disables FP contraction: fusing a rounded product into the
// following add breaks the error-free property.
struct df {
  float hi, lo;
};
__device__ __forceinline__ df df_make(float hi, float lo) { return df{hi, lo}; }
__device__ __forceinline__ df two_sum(float a, float b) {
#pragma clang fp contract(off)
  const float s = a + b;
  const float bb = s - a;
  const float e = (a - (s - bb)) + (b - bb);
  return df{s, e};
}
__device__ __forceinline__ df quick_two_sum(float a, float b) {
#pragma clang fp contract(off)
  const float s = a + b;
  return df{s, b - (s - a)};
}
__device__ __forceinline__ df two_prod(float a, float b) {
#pragma clang fp contract(off)
  const float p = a * b;
  return df{p, fmaf(a, b, -p)};
}
__device__ __forceinline__ df df_add(df x, df y) {
#pragma clang fp contract(off)
  df s = two_sum(x.hi, y.hi);
  return quick_two_sum(s.hi, s.lo + (x.lo + y.lo));
}
__device__ __forceinline__ df df_neg(df x) { return df{-x.hi, -x.lo}; }
__device__ __forceinline__ df df_sub(df x, df y) { return df_add(x, df_neg(y)); }
__device__ __forceinline__ df df_mul(df x, df y) {
#pragma clang fp contract(off)
  df p = two_prod(x.hi, y.hi);
  return quick_two_sum(p.hi, p.lo + (x.hi * y.lo + x.lo * y.hi));
}
__device__ __forceinline__ df df_sqrt(df x) {
#pragma clang fp contract(off)
  if (x.hi <= 0.0f) return df{0.0f, 0.0f};
  const float s = sqrtf(x.hi);
  df p = two_prod(s, s);
  const float corr = ((x.hi - p.hi) - p.lo + x.lo) / (2.0f * s);
  return quick_two_sum(s, corr);
}
__device__ __forceinline__ df df_div(df x, df y) {
#pragma clang fp contract(off)
  const float q1 = x.hi / y.hi;
  df r = df_sub(x, df_mul(df{q1, 0.0f}, y));
  const float q2 = (r.hi + r.lo) / y.hi;
  return quick_two_sum(q1, q2);
}

// ---------------------------------------------------------------- optics
// Soft probe aperture clip((semiangle - alpha)/denom + 0.5, 0, 1) with
// denom = sqrt((dx*ang_y)^2 + (dy*ang_x)^2)/|d|  (cuda/kernels/common.py:115-171).
//
// Normalising Gamma turns any non-zero aperture -- even 1e-7 -- into a
// full-weight contribution, so the outcome at the edge (edge == 0) and for
// A(Q-k) == A(Q+k) must be decided as exact arithmetic decides it.  Outside
// the soft band the value is exactly 0 or 1 (FP32 test with a safety
// margin); inside it the expression is evaluated in double-float from
// double-float coordinates and constants, then rounded to FP32.
struct Optics {
  float wavelength;
  float semiangle;   // rad
  float ang_y;       // rad per detector row pixel
  float ang_x;       // rad per detector column pixel
  float inner2;      // |d|^2 below which the aperture is exactly 1 (0 if unused)
  float outer2;      // |d|^2 above which the aperture is exactly 0 (with margin)
  df lam, semi, angy, angx;   // double-float constants
};

__device__ __forceinline__ float aperture_band(df dx, df dy, const Optics& o) {
#pragma clang fp contract(off)
  const df r2 = df_add(df_mul(dx, dx), df_mul(dy, dy));
  const df r = df_sqrt(r2);
  if (r.hi <= 1e-15f) return 1.0f;
  const df ty = df_mul(dx, o.angy), tx = df_mul(dy, o.angx);
  const df denom = df_div(df_sqrt(df_add(df_mul(ty, ty), df_mul(tx, tx))), r);
  if (denom.hi <= 1e-15f) return 1.0f;
  // edge = (semi - lam*r + denom/2) / denom
  const df t = df_add(df_sub(o.semi, df_mul(o.lam, r)), df{0.5f * denom.hi, 0.5f * denom.lo});
  const df edge = df_div(t, denom);
  const float e = edge.hi + edge.lo;
  return fminf(fmaxf(e, 0.0f), 1.0f);
}

// (dx, dy) = (qx - kx, qy - ky) or (qx + kx, qy + ky); *l are the low words.
template <bool PLUS>
__device__ __forceinline__ float aperture_at(float qx, float qxl, float qy, float qyl, float kx, float kxl,
                                             float ky, float kyl, const Optics& o) {
  const float dx = PLUS ? qx + kx : qx - kx;
  const float dy = PLUS ? qy + ky : qy - ky;
  const float r2 = dx * dx + dy * dy;
  if (r2 <= o.inner2) return 1.0f;
  if (r2 > o.outer2) return 0.0f;
  const df ddx = PLUS ? df_add(df{qx, qxl}, df{kx, kxl}) : df_sub(df{qx, qxl}, df{kx, kxl});
  const df ddy = PLUS ? df_add(df{qy, qyl}, df{ky, kyl}) : df_sub(df{qy, qyl}, df{ky, kyl});
  return aperture_band(ddx, ddy, o);
}

// ---------------------------------------------------------------- overlap
// For chi(u) = u^T M u (C10, C12, phi12 only) the reference overlap
//   Gamma = P(Q-k) P*(k) - P*(Q+k) P(k),  P(u) = A(u) exp(-i chi(u))
// factorises exactly:
//   Gamma = A(k) e^{i theta} [A(Q-k) e^{-i chi(Q)} - A(Q+k) e^{+i chi(Q)}],
//   theta = 2 Q^T M k = Qx*a_k + Qy*b_k.
// `eth` = e^{i theta}, `echi` = (cos chi(Q), sin chi(Q)).  The guard and
// normalisation follow cuda/kernels/common.py:227-230 on Gamma itself.
// Returns G * conj(Gamma^).
struct Coord {
  float x, xl, y, yl;   // double-float (hi, lo) of a 2-vector
};

__device__ __forceinline__ float2 trotter_term(float2 G, const Coord& q, const Coord& k, float ak, float2 eth,
                                               float2 echi, const Optics& o) {
  const float am = aperture_at<false>(q.x, q.xl, q.y, q.yl, k.x, k.xl, k.y, k.yl, o);
  const float ap = aperture_at<true>(q.x, q.xl, q.y, q.yl, k.x, k.xl, k.y, k.yl, o);
  // W = A_m (c - i s) - A_p (c + i s)
  const float wr = (am - ap) * echi.x;
  const float wi = -(am + ap) * echi.y;
  // Gamma = A_k * eth * W
  float gr = ak * (eth.x * wr - eth.y * wi);
  float gi = ak * (eth.x * wi + eth.y * wr);
  const float mag2 = gr * gr + gi * gi;
  const float scale = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
  gr *= scale;
  gi *= scale;
  return make_float2(G.x * gr + G.y * gi, G.y * gr - G.x * gi);
}

// General overlap from explicit chi values (any aberration order):
//   Gamma = A_m A_k e^{-i(chi_m - chi_k)} - A_p A_k e^{+i(chi_p - chi_k)}.
// Each phase difference is formed before its sincos so the large common
// parts cancel in FP32 before the trigonometric call.
__device__ __forceinline__ float2 trotter_term_general(float2 G, float am, float ap, float ak,
                                                       float dchi_m, float dchi_p) {
  float sm, cm, sp, cp;
  sincosf(dchi_m, &sm, &cm);
  sincosf(dchi_p, &sp, &cp);
  float gr = ak * (am * cm - ap * cp);
  float gi = ak * (-am * sm - ap * sp);
  const float mag2 = gr * gr + gi * gi;
  const float scale = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
  gr *= scale;
  gi *= scale;
  return make_float2(G.x * gr + G.y * gi, G.y * gr - G.x * gi);
}

}  // namespace qssb
