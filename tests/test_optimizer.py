"""Host fit drivers (no GPU): Nelder-Mead semantics of the reference and the
batched quasi-random search, on analytic objectives."""
from __future__ import annotations

import math

import numpy as np
import pytest

from quantem.gpu_amd.ssb.compute.hip import optimizer as opt


def test_nelder_mead_keeps_first_vertex_for_tied_losses():
    """``tests/test_ssb_batch_optuna.py:39-62``: a flat objective keeps x0."""
    flat = lambda pts: np.ones(len(pts), dtype=np.float32)
    x0 = np.asarray([-20.0, 22.0, 0.46])
    first = opt.nelder_mead(flat, x0, max_iter=4)
    second = opt.nelder_mead(flat, x0, max_iter=4)
    np.testing.assert_array_equal(first[0], x0)
    np.testing.assert_array_equal(second[0], first[0])
    assert first[1:] == second[1:]


def bowl(pts):
    pts = np.asarray(pts)
    return (pts[:, 0] - 123.0) ** 2 / 1e4 + (pts[:, 1] - 20.0) ** 2 / 100 + (pts[:, 2] - 0.2) ** 2


def test_nelder_mead_converges_and_counts_evaluations():
    calls = []

    def f(pts):
        calls.append(len(pts))
        return bowl(pts)

    x, loss, n = opt.nelder_mead(f, [100.0, 10.0, 0.0], xatol=1e-3, fatol=1e-12, max_iter=500)
    np.testing.assert_allclose(x, [123.0, 20.0, 0.2], atol=5e-3)
    assert loss < 1e-8
    assert calls[0] == 4 and n == sum(calls)
    with pytest.raises(ValueError, match="3 parameters"):
        opt.nelder_mead(f, [1.0, 2.0])


def test_batched_search_is_seeded_and_finds_the_basin():
    def f(params):
        return bowl([[p["C10_nm"], p["C12_nm"], math.radians(p["phi12_deg"])] for p in params])

    a = opt.batched_search(f, None, {}, 256, seed=7)
    b = opt.batched_search(f, None, {}, 256, seed=7)
    assert a[0] == b[0] and a[1] == b[1] and len(a[2]) == 256
    assert abs(a[0]["C10_nm"] - 123.0) < 40 and abs(a[0]["C12_nm"] - 20.0) < 8
    fixed = opt.batched_search(f, {"C10_nm": (-400.0, 400.0), "C12_nm": 20.0, "phi12_deg": 0.0}, {}, 32)
    assert set(fixed[0]) == {"C10_nm"}
    assert all(set(h["params"]) == {"C10_nm"} for h in fixed[2])
    with pytest.raises(ValueError):
        opt.batched_search(f, {"C10_nm": (5.0, -5.0)}, {}, 8)
