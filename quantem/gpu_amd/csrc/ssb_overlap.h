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

// G * conj(g)
__device__ __forceinline__ float2 mul_conj(float2 G, float2 g) {
  return make_float2(G.x * g.x + G.y * g.y, G.y * g.x - G.x * g.y);
}

}  // namespace qssb
