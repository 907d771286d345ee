"""The learned initializer's CPU oracle: plain-FP64 backward pinned by
central finite differences, bf16 rounding pinned by known values, and the
counter-based generator's statistics."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import learned_oracle as L


def _params(rng, scale=0.2):
    ps = []
    for cin, cout in L.CONV:
        ps.append((rng.standard_normal((cout, 9 * cin)) * scale / np.sqrt(9 * cin) * 3,
                   rng.standard_normal(cout) * 0.05))
    for fin, fout in L.FC:
        ps.append((rng.standard_normal((fout, fin)) / np.sqrt(fin), rng.standard_normal(fout) * 0.05))
    return ps


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e38, 1e-40], dtype=np.float32)
    # ties go to even: 1+2^-8 -> 1, 1+3*2^-8 -> 1+2^-6
    np.testing.assert_array_equal(L.bf16(x)[:5], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5])
    assert np.isfinite(L.bf16(x)[5])


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    params = _params(rng)
    X = rng.standard_normal((2, L.WIN, L.WIN, 16))
    t = rng.standard_normal((2, 3))
    pred, cache = L.forward(params, X, quantize=False)
    _, dpred = L.mse(pred, t)
    grads = L.backward(params, cache, dpred, quantize=False)

    def loss_at(ps):
        return L.mse(L.forward(ps, X, quantize=False)[0], t)[0]

    h = 1e-6
    for li in (0, 2, 3, 4, 5):
        w, b = params[li]
        for (i, j) in ((0, 0), (1, 5), (w.shape[0] - 1, w.shape[1] - 1)):
            wp = w.copy(); wp[i, j] += h
            wm = w.copy(); wm[i, j] -= h
            pp = list(params); pp[li] = (wp, b)
            pm = list(params); pm[li] = (wm, b)
            num = (loss_at(pp) - loss_at(pm)) / (2 * h)
            assert grads[li][0][i, j] == pytest.approx(num, rel=1e-5, abs=1e-9)
        bp = b.copy(); bp[1] += h
        bm = b.copy(); bm[1] -= h
        pp = list(params); pp[li] = (w, bp)
        pm = list(params); pm[li] = (w, bm)
        num = (loss_at(pp) - loss_at(pm)) / (2 * h)
        assert grads[li][1][1] == pytest.approx(num, rel=1e-5, abs=1e-9)


def test_im2col_col2im_are_adjoint():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 8, 8, 3))
    D = rng.standard_normal((2 * 4 * 4, 27))
    assert np.sum(L.im2col(X) * D) == pytest.approx(np.sum(X * L.col2im(D, 2, 8, 8, 3)), rel=1e-12)


def test_generator_statistics():
    u = L.uniform01(7, 3, np.arange(200000, dtype=np.uint64))
    assert 0.0 < u.min() and u.max() < 1.0 and abs(u.mean() - 0.5) < 0.003
    zr, zi = L.normal2(7, 3, np.arange(100000, dtype=np.uint64))
    assert abs(zr.std() - 1) < 0.01 and abs(zi.mean()) < 0.01 and abs(np.corrcoef(zr, zi)[0, 1]) < 0.01
