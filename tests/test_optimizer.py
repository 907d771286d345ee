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


def _quadratic(A, centre, floor=0.0):
    A = np.asarray(A, dtype=np.float64)
    c = np.asarray(centre, dtype=np.float64)

    def evaluate(points):
        P = np.atleast_2d(points) - c
        return floor + 0.5 * np.einsum("ki,ij,kj->k", P, A, P)
    return evaluate


def test_stencil_model_recovers_gradient_and_hessian_of_a_quadratic():
    A = np.array([[4.0, 1.9, 0.2], [1.9, 1.0, 0.1], [0.2, 0.1, 2.0]])
    c = np.array([0.3, -0.7, 1.2])
    x = np.array([1.0, 0.5, -0.25])
    h = np.array([0.1, 0.2, 0.05])
    unit = opt._stencil(3)
    assert unit.shape == (10, 3)                      # centre, 6 axis points, 3 pairs
    g, H = opt._stencil_model(_quadratic(A, c)(x + unit * h), h)
    np.testing.assert_allclose(H, A, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(g, A @ (x - c), rtol=1e-9, atol=1e-9)


def test_newton_refine_lands_on_a_quadratic_minimum_in_few_batches():
    A = np.array([[4.0, 1.9, 0.2], [1.9, 1.0, 0.1], [0.2, 0.1, 2.0]])
    c = np.array([0.3, -0.7, 1.2])
    calls = []

    def evaluate(points):
        calls.append(len(points))
        return _quadratic(A, c, 3.0)(points)
    x, fx, n = opt.newton_refine(evaluate, np.ones(3), [0.1, 0.2, 0.05], xtol=1e-8, hmin=1e-9)
    np.testing.assert_allclose(x, c, atol=1e-9)
    assert fx == pytest.approx(3.0, abs=1e-15)
    assert all(k == 10 for k in calls) and n == sum(calls) and len(calls) <= 4


def test_newton_refine_follows_a_badly_scaled_valley_through_noise():
    # curvatures of the exact SSB loss in (C10 nm, C30, rotation deg) near an
    # optimum: a C10/C30 valley with condition number ~1e2 in stencil units,
    # plus 3e-8 of evaluation noise
    H = np.array([[8.4e-6, 1.56e-9, 0.0], [1.56e-9, 3.55e-13, 0.0], [0.0, 0.0, 4.8e-4]])
    c = np.array([283.757, 58581.2, -2.6422])
    rng = np.random.default_rng(0)
    base = _quadratic(H, c, 0.007944)

    def evaluate(points):
        return base(points) + 3e-8 * rng.standard_normal(len(np.atleast_2d(points)))
    x, fx, n = opt.newton_refine(evaluate, [300.645, -20976.9, -2.6386], [10.0, 1e4, 0.25],
                                 xtol=[0.05, 50.0, 2e-3], hmin=[1.0, 2500.0, 0.1], ftol=1e-8)
    assert abs(x[0] - c[0]) < 0.5 and abs(x[1] - c[1]) < 3000.0 and abs(x[2] - c[2]) < 0.02
    assert n <= 80


def test_newton_refine_respects_bounds_and_validates_steps():
    x, fx, _ = opt.newton_refine(_quadratic(np.eye(2), [2.0, -3.0]), np.zeros(2), 0.25, max_iter=10,
                                 lo=np.array([-1.0, -1.0]), hi=np.array([1.0, 1.0]))
    np.testing.assert_allclose(x, [1.0, -1.0], atol=1e-6)
    with pytest.raises(ValueError):
        opt.newton_refine(_quadratic(np.eye(2), [0.0, 0.0]), np.zeros(2), [0.1, 0.0])


def test_newton_refine_stops_on_a_small_predicted_decrease():
    A = np.diag([2.0, 1.0, 4.0])
    calls = []

    def evaluate(points):
        calls.append(len(points))
        return _quadratic(A, [0.1, -0.2, 0.3], 1.0)(points)
    # the first step (inside the trust box of 4 stencil widths) solves the
    # quadratic; the second batch, at the minimum, predicts ~0 decrease,
    # below ftol, and stops.  The polish needs 2 stencils' worth of points
    # near the best one; the first batch lies farther out, so it is skipped.
    x, fx, n = opt.newton_refine(evaluate, np.zeros(3), 0.1, max_iter=8, ftol=1e-12, polish=True)
    np.testing.assert_allclose(x, [0.1, -0.2, 0.3], atol=1e-9)
    assert calls == [10, 10] and n == 20


def test_newton_refine_keeps_a_symmetric_stencil_at_a_bound():
    # minimum just inside the upper bound of x0 (rotation clipped to +-5 deg):
    # a stencil of half-width 0.25 around 4.95 would cross 5.  Every point
    # stays inside the box and every batch is a central stencil (the axis
    # pairs straddle their centre at equal distances), so the model is the
    # quadratic itself and one step lands on the minimum.
    A = np.array([[2.0, 0.3], [0.3, 1.0]])
    c = np.array([4.9, 1.0])
    batches = []

    def evaluate(points):
        batches.append(np.array(points, copy=True))
        return _quadratic(A, c, 0.5)(points)
    lo, hi = np.array([-5.0, -5.0]), np.array([5.0, 5.0])
    x, fx, _ = opt.newton_refine(evaluate, [4.95, 0.0], [0.25, 0.25], max_iter=4, xtol=1e-9, hmin=1e-6,
                                 lo=lo, hi=hi)
    np.testing.assert_allclose(x, c, atol=1e-9)
    for pts in batches:
        assert np.all(pts >= lo) and np.all(pts <= hi)
        if len(pts) == len(opt._stencil(2)):
            centre, plus, minus = pts[0], pts[1:3], pts[3:5]
            np.testing.assert_allclose(plus - centre, centre - minus, atol=1e-12)


def test_newton_refine_reaches_a_minimum_outside_the_box_at_the_bound():
    # constrained minimum on the bound: the result is the bound itself
    x, fx, _ = opt.newton_refine(_quadratic(np.eye(1), [6.0]), [4.0], [0.5], max_iter=10, lo=np.array([-5.0]),
                                 hi=np.array([5.0]))
    assert x[0] == pytest.approx(5.0, abs=1e-9)


def _rough(A, c, step):
    """Quadratic bowl 0.5 (x - c)^T A (x - c) plus a staircase of height
    ``step`` that flips sign on a fine pseudo-random grid: a loss whose lowest
    sample near the minimum is decided by the staircase."""
    A = np.asarray(A, float)
    c = np.asarray(c, float)

    def f(points):
        x = np.atleast_2d(np.asarray(points, float))
        d = x - c
        smooth = 0.5 * np.einsum("ni,ij,nj->n", d, A, d)
        cells = np.floor(x / 0.013).astype(np.int64)
        h = (cells[:, 0] * 73856093) ^ (cells[:, -1] * 19349663)
        return smooth + step * np.where(h % 2 == 0, 1.0, -1.0)
    return f


def test_model_polish_averages_a_rough_loss_that_the_lowest_sample_does_not():
    A = [[2.0, 0.4, 0.0], [0.4, 1.0, 0.2], [0.0, 0.2, 3.0]]
    c = np.array([0.31, -0.17, 0.08])
    f = _rough(A, c, 2e-3)
    x0 = c + np.array([0.6, -0.5, 0.4])
    seen = []

    def evaluate(points):
        v = f(points)
        seen.extend(zip(np.atleast_2d(points), v))
        return v
    x, n = opt.model_polish(evaluate, x0, [0.5, 0.5, 0.5], rounds=2, batch=32, seed=3)
    assert n == 64
    err_model = float(np.linalg.norm(x - c))
    lowest = min(seen, key=lambda t: t[1])[0]
    assert err_model < 0.03
    assert err_model < float(np.linalg.norm(lowest - c))


def test_model_polish_recovers_an_exact_quadratic_and_respects_bounds():
    A = np.diag([1.0, 4.0])
    x, _ = opt.model_polish(_quadratic(A, [0.2, -0.1]), [1.0, 1.0], [1.0, 1.0], rounds=2, batch=8)
    np.testing.assert_allclose(x, [0.2, -0.1], atol=1e-9)
    lo, hi = np.array([-1.0, 0.0]), np.array([1.0, 1.0])
    x, _ = opt.model_polish(_quadratic(A, [0.2, -0.5]), [0.5, 0.5], [0.5, 0.5], rounds=3, batch=8, lo=lo, hi=hi)
    assert np.all(x >= lo) and np.all(x <= hi)
    assert x[1] == pytest.approx(0.0, abs=1e-9)
    with pytest.raises(ValueError, match="batch"):
        opt.model_polish(_quadratic(A, [0.0, 0.0]), [0.0, 0.0], [1.0, 1.0], batch=5)
