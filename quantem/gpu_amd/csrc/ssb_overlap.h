// Per-point pieces of the SSB overlap shared by the object path and the
// exact per-BF loss path: the factorised Gamma (see ssb_device.h), the
// isotropic double-float aperture and the quadratic chi.
#pragma once

#include "ssb_internal.h"

namespace qssb {

// Balanced split of [0, nbf) into `groups` contiguous ranges.
__device__ __forceinline__ void group_range(int g, int groups, int nbf, int& b0, int& b1) {
  b0 = (int)(((long long)g * nbf) / groups);
  b1 = (int)(((long long)(g + 1) * nbf) / groups);
}

__device__ __forceinline__ float chi_quad(const QuadChi& m, float x, float y) {
  return m.mxx * x * x + 2.0f * m.mxy * x * y + m.myy * y * y;
}

// Aperture from d = Rout^2 - |u|^2 (double-float), where Rout is the radius
// at which the soft edge reaches 0.  For isotropic detector sampling the
// reference edge (semi - lambda|u|)/ang + 1/2 equals (Rout - |u|)*lambda/ang
// = d * (lambda/ang) / (Rout + |u|): the cancellation sits entirely in d,
// which is exact to ~1e-13 here, so ties (d == 0, A(Q-k) == A(Q+k)) resolve
// as exact arithmetic resolves them.
__device__ __forceinline__ float aperture_from_d(df d, const ApIso& a) {
  const float dh = d.hi + d.lo;
  if (dh >= a.d_in) return 1.0f;
  if (dh <= a.d_out) return 0.0f;
  const float r = sqrtf(fmaxf(a.rout2 - dh, 0.0f));
  const float e = dh * a.scale / (a.rout + r);
  return fminf(fmaxf(e, 0.0f), 1.0f);
}

// aperture_from_d(a + b) for d given as two double-float terms: the float32
// sum of their high words is within a few 1e-6 of d for the magnitudes here
// (|a|, |b| < 32), so outside the soft band by more than 1e-4 the value is
// exactly 1 or 0 without the double-float addition; inside, the same
// double-float evaluation as aperture_from_d.
__device__ __forceinline__ float aperture_sum(const df& a, const df& b, const ApIso& ap) {
  constexpr float kMargin = 1e-4f;
  const float est = a.hi + b.hi;
  if (est >= ap.d_in + kMargin) return 1.0f;
  if (est <= ap.d_out - kMargin) return 0.0f;
  return aperture_from_d(df_add(a, b), ap);
}

// The aperture from a float32 estimate of d when it lies beyond the 1e-4
// margin of the soft band (the value aperture_sum gives there); false when
// the estimate is too close to the band to decide.
__device__ __forceinline__ bool aperture_decided(float est, const ApIso& ap, float& a) {
  constexpr float kMargin = 1e-4f;
  a = (est >= ap.d_in + kMargin) ? 1.0f : 0.0f;
  return est >= ap.d_in + kMargin || est <= ap.d_out - kMargin;
}

// The value unchanged, through an empty volatile asm: an expression built on
// it cannot be hoisted out of the branch that holds it.
__device__ __forceinline__ df df_in_branch(df x) {
  asm volatile("" : "+v"(x.hi), "+v"(x.lo));
  return x;
}

// Float32 estimates of the isotropic row terms of d(Q -/+ k),
// -qx^2 + 2 qx kx and -qx^2 - 2 qx kx (kx2 = 2 kx): within ~1e-5 of the
// double-float values for terms below 32 in magnitude, well inside the 1e-4
// margin that aperture_decided leaves.
__device__ __forceinline__ float row_estimate_minus(df qx, df kx2, df q2x) { return fmaf(qx.hi, kx2.hi, -q2x.hi); }
__device__ __forceinline__ float row_estimate_plus(df qx, df kx2, df q2x) { return -fmaf(qx.hi, kx2.hi, q2x.hi); }

// Isotropic aperture pair A(Q - k), A(Q + k) from d = row + col: decided
// from the estimates uq, vq of the row terms wherever d lies beyond the soft
// band's margin (exactly the value aperture_sum gives), otherwise from the
// double-float row terms, which are formed inside that branch only -- the
// few points near the soft edge pay for them, the rest do not.
__device__ __forceinline__ void iso_aperture_pair(float uq, float vq, df qx, df kx2, df q2x, const df& colA,
                                                  const df& colB, const ApIso& ap, float& am, float& app) {
  const bool da = aperture_decided(uq + colA.hi, ap, am);
  const bool dp = aperture_decided(vq + colB.hi, ap, app);
  if (!(da && dp)) {
    const df u = df_mul(df_in_branch(qx), kx2);
    if (!da) am = aperture_from_d(df_add(df_sub(u, q2x), colA), ap);
    if (!dp) app = aperture_from_d(df_add(df_neg(df_add(q2x, u)), colB), ap);
  }
}

// Aperture pair A(Q - k), A(Q + k).  ISO: from d = (Rout^2 - |k|^2) - |Q|^2
// +/- 2 Q.k; otherwise the general double-float soft-edge expression.
template <bool ISO>
__device__ __forceinline__ void aperture_pair(df qx, df qy, df kx, df ky, df cb, df q2, const ApIso& ap,
                                              const Optics& o, float& am, float& app) {
  if constexpr (ISO) {
    const df qk2 = df_mul(df_add(df_mul(qx, kx), df_mul(qy, ky)), df{2.0f, 0.0f});
    const df base = df_sub(cb, q2);
    am = aperture_from_d(df_add(base, qk2), ap);
    app = aperture_from_d(df_sub(base, qk2), ap);
  } else {
    am = aperture_at<false>(qx.hi, qx.lo, qy.hi, qy.lo, kx.hi, kx.lo, ky.hi, ky.lo, o);
    app = aperture_at<true>(qx.hi, qx.lo, qy.hi, qy.lo, kx.hi, kx.lo, ky.hi, ky.lo, o);
  }
}


// Normalised overlap Gamma^ = Gamma / |Gamma| of one (b, Q) from its
// factorised parts, Gamma = A(k) e^{i theta} [A(Q-k) e^{-i chi(Q)} - A(Q+k) e^{i chi(Q)}],
// with the guard |Gamma|^2 > 1e-16 ? rsqrt : 1e8 (cuda/kernels/common.py:227-230).
__device__ __forceinline__ float2 gamma_hat(float ak, float2 eth, float am, float app, float2 echi) {
  const float wr = (am - app) * echi.x;
  const float wi = -(am + app) * echi.y;
  float gr = ak * (eth.x * wr - eth.y * wi);
  float gi = ak * (eth.x * wi + eth.y * wr);
  const float mag2 = gr * gr + gi * gi;
  const float sc = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
  return make_float2(gr * sc, gi * sc);
}

// chi(u) split into its even part E and odd part O (ChiPoly).
__device__ __forceinline__ void chi_parts(const ChiPoly& cp, float ux, float uy, float& E, float& O) {
  const float r2 = ux * ux + uy * uy;
  float zr[7], zi[7];
  zr[0] = 1.0f;
  zi[0] = 0.0f;
#pragma unroll
  for (int m = 1; m < 7; ++m) {
    zr[m] = zr[m - 1] * ux - zi[m - 1] * uy;
    zi[m] = zr[m - 1] * uy + zi[m - 1] * ux;
  }
  const float r4 = r2 * r2, r6 = r4 * r2;
  auto t = [&](int i, int m) { return cp.ca[i] * zr[m] + cp.sa[i] * zi[m]; };
  E = cp.K[0] * r2 + cp.K[1] * t(1, 2) + cp.K[4] * r4 + cp.K[5] * r2 * t(5, 2) + cp.K[6] * t(6, 4) +
      cp.K[10] * r6 + cp.K[11] * r4 * t(11, 2) + cp.K[12] * r2 * t(12, 4) + cp.K[13] * t(13, 6);
  O = cp.K[2] * r2 * t(2, 1) + cp.K[3] * t(3, 3) + cp.K[7] * r4 * t(7, 1) + cp.K[8] * r2 * t(8, 3) +
      cp.K[9] * t(9, 5);
}

// Gamma^ from the phases: Gamma = A(k) [A(Q-k) e^{-i dm} - A(Q+k) e^{i dp}],
// dm = chi(Q-k) - chi(k), dp = chi(Q+k) - chi(k); same guard as gamma_hat.
__device__ __forceinline__ float2 gamma_hat_chi(float ak, float am, float app, float dm, float dp) {
  if (am == 0.0f && app == 0.0f) return make_float2(0.0f, 0.0f);
  float sm, cm, sp, cpp;
  sincosf(dm, &sm, &cm);
  sincosf(dp, &sp, &cpp);
  float gr = ak * (am * cm - app * cpp);
  float gi = ak * (-am * sm - app * sp);
  const float mag2 = gr * gr + gi * gi;
  const float sc = (mag2 > 1e-16f) ? rsqrtf(mag2) : 1e8f;
  return make_float2(gr * sc, gi * sc);
}

// G * conj(g)
__device__ __forceinline__ float2 mul_conj(float2 G, float2 g) {
  return make_float2(G.x * g.x + G.y * g.y, G.y * g.x - G.x * g.y);
}

}  // namespace qssb
