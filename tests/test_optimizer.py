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


def _counting(fn):
    calls = []

    def evaluate(pts):
        pts = np.asarray(pts, dtype=np.float64)
        calls.append(len(pts))
        return np.array([fn(p) for p in pts])
    return evaluate, calls


def test_quadratic_refine_solves_a_correlated_quadratic_in_one_round():
    """A full quadratic with a strongly correlated valley (the C10/C30
    trade-off shape) is fitted exactly, so one round lands on its minimiser."""
    H = np.array([[2.0, 1.9, 0.1], [1.9, 2.0, 0.0], [0.1, 0.0, 0.5]])
    xs = np.array([0.3, -0.2, 0.1])
    f = lambda x: 0.01 + 0.5 * (x - xs) @ H @ (x - xs)
    evaluate, calls = _counting(f)
    x, fx, n = opt.quadratic_refine(evaluate, np.zeros(3), radius=1.0, rounds=1, batch=16)
    assert np.allclose(x, xs, atol=1e-9) and fx == pytest.approx(0.01, abs=1e-12)
    assert n == 17 and calls == [16, 1]      # the batch, then the model's minimiser


def test_quadratic_refine_converges_on_a_non_quadratic_bowl():
    v = np.array([1.0, -1.0, 0.5]) / 1.5
    xs = np.array([0.7, -0.4, 0.25])
    f = lambda x: float(np.sum((x - xs) ** 2 * [1.0, 3.0, 0.3]) + 2.0 * ((x - xs) @ v) ** 4
                        + 0.3 * ((x - xs) @ v) ** 3)
    evaluate, calls = _counting(f)
    x, fx, n = opt.quadratic_refine(evaluate, np.zeros(3), radius=np.array([1.0, 1.0, 1.0]), rounds=7, batch=16)
    assert np.max(np.abs(x - xs)) < 1e-4
    assert n == sum(calls) and 7 * 16 < n <= 7 * 17 and calls[0] == 16
    x2, fx2, _ = opt.quadratic_refine(evaluate, np.zeros(3), radius=1.0, rounds=7, batch=16)
    assert np.array_equal(x, x2) and fx == fx2   # seeded: deterministic


def test_quadratic_refine_falls_back_to_the_best_point_and_respects_bounds():
    f = lambda x: -float(x @ x)      # concave: no positive-definite model
    evaluate, _ = _counting(f)
    x, fx, _ = opt.quadratic_refine(evaluate, np.zeros(2), radius=0.5, rounds=2, batch=8,
                                    lo=[-0.6, -0.6], hi=[0.6, 0.6])
    assert fx <= -0.25 and np.all(np.abs(x) <= 0.6)
    with pytest.raises(ValueError, match="batch"):
        opt.quadratic_refine(evaluate, np.zeros(3), radius=1.0, batch=8)
