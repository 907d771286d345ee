"""CPU oracle of the learned initializer's network and training features.

TEST INFRASTRUCTURE ONLY: ``tests/`` use it to check the GPU kernels of
``quantem/gpu_amd/csrc/ssb_learn.hip``; nothing in the package imports it.
The network has no counterpart in the reference (it only seeds the batched
search), so this file restates the library's own design:

* ``forward``/``backward`` in FP64, with bf16 rounding (round to nearest
  even) applied at exactly the points where the GPU rounds: GEMM operands
  (im2col columns, activations entering a fully connected layer, weights,
  the ReLU-masked output gradients); accumulation is exact here, fp32 there.
  With ``quantize=False`` the same code is plain FP64 and is pinned against
  central finite differences.
* ``features_synth``: the weak-phase forward-model features, including the
  counter-based hash generator, so GPU features can be compared value by
  value.
"""
from __future__ import annotations

import math

import numpy as np

CONV = ((16, 32), (32, 64), (64, 64), (64, 64))
FC = ((256, 256), (256, 3))
WIN = 32
NQ = 8


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def im2col(X: np.ndarray) -> np.ndarray:
    """NHWC -> [(b, oy, ox), (ky, kx, c)] for 3x3, stride 2, pad 1."""
    B, H, W, C = X.shape
    Ho, Wo = H // 2, W // 2
    P = np.zeros((B, H + 2, W + 2, C))
    P[:, 1:H + 1, 1:W + 1] = X
    cols = np.empty((B, Ho, Wo, 3, 3, C))
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky, kx] = P[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2]
    return cols.reshape(B * Ho * Wo, 9 * C)


def col2im(dcols: np.ndarray, B: int, H: int, W: int, C: int) -> np.ndarray:
    Ho, Wo = H // 2, W // 2
    d = dcols.reshape(B, Ho, Wo, 3, 3, C)
    P = np.zeros((B, H + 2, W + 2, C))
    for ky in range(3):
        for kx in range(3):
            P[:, ky:ky + 2 * Ho:2, kx:kx + 2 * Wo:2] += d[:, :, :, ky, kx]
    return P[:, 1:H + 1, 1:W + 1]


def forward(params, X, quantize: bool = True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    cache = []
    h = np.asarray(X, dtype=np.float64)
    for li, (cin, cout) in enumerate(CONV):
        w, b = params[li]
        B, H = h.shape[0], h.shape[1]
        cols = q(im2col(h))
        y = np.maximum(cols @ q(w).T + b, 0.0)
        cache.append((h.shape, cols, y))
        h = y.reshape(B, H // 2, H // 2, cout)
    flat = h.reshape(h.shape[0], -1)
    for fi in range(len(FC)):
        w, b = params[len(CONV) + fi]
        xh = q(flat)
        y = xh @ q(w).T + b
        if fi == 0:
            y = np.maximum(y, 0.0)
        cache.append((flat.shape, xh, y))
        flat = y
    return flat, cache


def mse(pred, target):
    d = pred - target
    return float(np.mean(d * d)), 2.0 * d / d.size


def backward(params, cache, dpred, quantize: bool = True):
    q = bf16 if quantize else (lambda a: np.asarray(a, dtype=np.float64))
    grads = [None] * len(params)
    dy = dpred
    for fi in reversed(range(len(FC))):
        li = len(CONV) + fi
        _, xh, y = cache[li]
        dz = dy * (y > 0) if fi == 0 else dy
        dzh = q(dz)
        grads[li] = (dzh.T @ xh, dz.sum(axis=0))
        dy = dzh @ q(params[li][0])
    for li in reversed(range(len(CONV))):
        shape, cols, y = cache[li]
        dz = dy.reshape(y.shape) * (y > 0)
        dzh = q(dz)
        grads[li] = (dzh.T @ cols, dz.sum(axis=0))
        if li > 0:
            B, H, W, C = shape
            dy = col2im(dzh @ q(params[li][0]), B, H, W, C)
    return grads


# --------------------------------------------------------------- features
def _hash32(x):
    x = np.asarray(x, dtype=np.uint64) & 0xFFFFFFFF
    x ^= x >> 16
    x = (x * 0x7FEB352D) & 0xFFFFFFFF
    x ^= x >> 15
    x = (x * 0x846CA68B) & 0xFFFFFFFF
    x ^= x >> 16
    return x


def uniform01(seed, a, b):
    inner = _hash32((np.uint64(b) + 0x632BE5AB) & 0xFFFFFFFF)
    mid = _hash32(((np.uint64(a) * 0x9E3779B9) & 0xFFFFFFFF) ^ inner)
    h = _hash32(np.uint64(seed) ^ mid)
    return ((h >> 8).astype(np.float64) + 0.5) / 16777216.0


def normal2(seed, a, b):
    u1 = uniform01(seed, a, 2 * np.asarray(b, dtype=np.uint64))
    u2 = uniform01(seed, a, 2 * np.asarray(b, dtype=np.uint64) + 1)
    r = np.sqrt(-2.0 * np.log(u1))
    return r * np.cos(2 * math.pi * u2), r * np.sin(2 * math.pi * u2)


def _aperture(ux, uy, semi, lam, ang):
    return np.clip((semi - lam * np.sqrt(ux * ux + uy * uy)) / ang + 0.5, 0.0, 1.0)


def _chi(ux, uy, c10, c30, lam):
    r2 = ux * ux + uy * uy
    return math.pi * lam * r2 * c10 + 0.5 * math.pi * lam ** 3 * r2 * r2 * c30


def features_synth(theta, snr, recip, semi, ang, lam, q, seed):
    """[S, 32, 32, 16] features of the forward model (same generator as the GPU)."""
    S = len(theta)
    out = np.zeros((S, WIN, WIN, 2 * NQ))
    iy, ix = np.mgrid[:WIN, :WIN]
    p = (iy * WIN + ix).astype(np.uint64)
    for s in range(S):
        c10, c30, rot = (float(v) for v in theta[s])
        oy = uniform01(seed ^ 0x27D4EB2F, s, 1) - 0.5
        ox = uniform01(seed ^ 0x27D4EB2F, s, 2) - 0.5
        kxa = (iy - WIN // 2 + oy) * recip
        kya = (ix - WIN // 2 + ox) * recip
        cr, sr = math.cos(-rot), math.sin(-rot)
        kx = kxa * cr + kya * sr
        ky = -kxa * sr + kya * cr
        ak = _aperture(kx, ky, semi, lam, ang)
        chik = _chi(kx, ky, c10, c30, lam)
        for qi in range(NQ):
            qx, qy = float(q[qi][0]), float(q[qi][1])
            phr, phi = normal2(seed, s, 1000 + qi)
            am = _aperture(qx - kx, qy - ky, semi, lam, ang)
            ap = _aperture(qx + kx, qy + ky, semi, lam, ang)
            dm = _chi(qx - kx, qy - ky, c10, c30, lam) - chik
            dp = _chi(qx + kx, qy + ky, c10, c30, lam) - chik
            gr = ak * (am * np.cos(dm) - ap * np.cos(dp))
            gi = ak * (-am * np.sin(dm) - ap * np.sin(dp))
            nr, ni = normal2(seed ^ 0x5BD1E995, s, qi * WIN * WIN + p)
            vr = gr * phr - gi * phi + snr[s] * nr
            vi = gr * phi + gi * phr + snr[s] * ni
            inside = ak > 0
            vr = np.where(inside, vr, 0.0)
            vi = np.where(inside, vi, 0.0)
            mag = np.sqrt(vr * vr + vi * vi)
            mx = mag.max()
            cr_, ci_ = vr[WIN // 2, WIN // 2], vi[WIN // 2, WIN // 2]
            out[s, :, :, 2 * qi] = mag / mx if mx > 0 else 0.0
            ang_ = np.arctan2(vi * cr_ - vr * ci_, vr * cr_ + vi * ci_) / math.pi
            out[s, :, :, 2 * qi + 1] = np.where(mag > 0, ang_, 0.0)
    return out
