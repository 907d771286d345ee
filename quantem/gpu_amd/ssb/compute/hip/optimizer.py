"""Host drivers of the aberration fit: a batched quasi-random search and the
Nelder-Mead refinement.

Both take ``evaluate(points) -> losses``: ``points`` is a float64 array
``[K, 3]`` of (C10, C12, phi12) and the callee evaluates all K on the GPU
with one synchronisation (the losses stay on the device until the batch is
done).  Nothing here computes a loss.

* ``nelder_mead`` follows ``batch_nelder_mead`` (``cuda/optimizer.py:237-396``):
  initial steps max(5% |x_i|, 2.5e-4), coefficients (alpha, gamma, rho, sigma)
  = (1, 2, 0.5, 0.5), vertices ordered by a stable sort so ties keep the
  earlier vertex, convergence when the simplex spread is below ``xatol`` and
  the loss spread below ``fatol``, the 4 initial vertices and the 3 shrunk
  vertices evaluated as batches.
* ``quadratic_refine`` is the batched local refinement: each round
  evaluates one batch (the centre, the axis points and a Sobol fill of a
  box), fits a full quadratic model by least squares and moves to its
  minimiser inside the box; it needs a few batches where Nelder-Mead needs
  100+ sequential evaluations.  It has no reference counterpart.
* ``batched_search`` replaces the reference's Optuna TPE loop
  (``cuda/optimizer.py:59-230``), whose trial sequence cannot be reproduced
  offline: a scrambled Sobol design over the search box (seeded), evaluated
  in batches, then a second Sobol design in a box shrunk around the best
  point.  Parameter names and the trial-history format follow the
  reference (``C10_nm``, ``C12_nm``, ``phi12_deg``).
"""
from __future__ import annotations

import math
from typing import Callable

import numpy as np

Evaluate = Callable[[np.ndarray], np.ndarray]

PARAM_KEYS = ("C10_nm", "C12_nm", "phi12_deg")
DEFAULT_RANGES = {"C10_nm": (-400.0, 400.0), "C12_nm": (0.0, 100.0), "phi12_deg": (-90.0, 90.0)}


def nelder_mead(evaluate: Evaluate, x0, xatol: float = 0.1, fatol: float = 1e-8, max_iter: int = 300,
                flat_fatol: float | None = None) -> tuple[np.ndarray, float, int]:
    """Minimise over (C10, C12, phi12) from ``x0``; returns (best_x, best_loss, n_evals)."""
    x0 = np.asarray(x0, dtype=np.float64)
    n = x0.size
    if n != 3:
        raise ValueError("Expected 3 parameters: C10, C12, phi12")
    simplex = np.empty((n + 1, n), dtype=np.float64)
    simplex[0] = x0
    for i in range(n):
        simplex[i + 1] = x0
        simplex[i + 1, i] += max(abs(x0[i]) * 0.05, 0.00025)
    f = np.asarray(evaluate(simplex), dtype=np.float64).copy()
    n_evals = n + 1
    if flat_fatol is not None and np.ptp(f) <= float(flat_fatol):
        return simplex[0], float(f[0]), n_evals
    alpha, gamma, rho, sigma = 1.0, 2.0, 0.5, 0.5

    def one(x) -> float:
        return float(np.asarray(evaluate(x[None, :]), dtype=np.float64)[0])

    for _ in range(max_iter):
        order = np.argsort(f, kind="stable")
        simplex = simplex[order]
        f = f[order]
        if np.max(np.abs(simplex[-1] - simplex[0])) < xatol and abs(f[-1] - f[0]) < fatol:
            break
        centroid = simplex[:-1].mean(axis=0)
        x_r = centroid + alpha * (centroid - simplex[-1])
        f_r = one(x_r)
        n_evals += 1
        if f[0] <= f_r < f[-2]:
            simplex[-1], f[-1] = x_r, f_r
            continue
        if f_r < f[0]:
            x_e = centroid + gamma * (x_r - centroid)
            f_e = one(x_e)
            n_evals += 1
            if f_e < f_r:
                simplex[-1], f[-1] = x_e, f_e
            else:
                simplex[-1], f[-1] = x_r, f_r
            continue
        if f_r < f[-1]:
            x_c = centroid + rho * (x_r - centroid)
            f_c = one(x_c)
            n_evals += 1
            if f_c <= f_r:
                simplex[-1], f[-1] = x_c, f_c
                continue
        else:
            x_c = centroid - rho * (centroid - simplex[-1])
            f_c = one(x_c)
            n_evals += 1
            if f_c < f[-1]:
                simplex[-1], f[-1] = x_c, f_c
                continue
        for i in range(1, n + 1):
            simplex[i] = simplex[0] + sigma * (simplex[i] - simplex[0])
        f[1:] = np.asarray(evaluate(simplex[1:]), dtype=np.float64)
        n_evals += n
    best = int(np.argsort(f, kind="stable")[0])
    return simplex[best], float(f[best]), n_evals


def _split_specs(aberrations: dict, defaults: dict, keys):
    """Free (key, lo, hi) ranges and fixed values, in the search's units."""
    free = []
    fixed = {}
    for key in keys:
        spec = aberrations.get(key, defaults.get(key, 0.0))
        if isinstance(spec, tuple):
            lo, hi = float(spec[0]), float(spec[1])
            if not hi >= lo:
                raise ValueError(f"search range for {key} must be (low, high); got {spec!r}")
            free.append((key, lo, hi))
        else:
            fixed[key] = float(spec)
    return free, fixed


def batched_search(evaluate_params: Callable[[list], np.ndarray], aberrations: dict | None, defaults: dict,
                   n_trials: int, seed: int = 42, batch: int = 64, refine_fraction: float = 0.25,
                   keys=PARAM_KEYS):
    """Quasi-random search over ``keys``; returns (best_params, best_value, history).

    ``evaluate_params(list_of_dicts)`` returns the losses of a batch of full
    parameter dicts (every key of ``keys``).  ``defaults`` gives the value of
    keys absent from ``aberrations``.
    """
    from scipy.stats import qmc

    if n_trials < 1:
        raise ValueError("n_trials must be positive")
    aberrations = dict(DEFAULT_RANGES if aberrations is None else aberrations)
    free, fixed = _split_specs(aberrations, defaults, keys)
    history: list[dict] = []
    full_params: list[dict] = []

    def run(box, count, rng_seed):
        if count <= 0:
            return
        if free:
            sampler = qmc.Sobol(d=len(free), scramble=True, seed=rng_seed)
            m = int(math.ceil(math.log2(max(count, 1))))
            unit = sampler.random_base2(m)[:count]
        else:
            unit = np.zeros((count, 0))
        params = []
        for u in unit:
            p = dict(fixed)
            for (key, _, _), (lo, hi), v in zip(free, box, u):
                p[key] = lo + (hi - lo) * float(v)
            params.append(p)
        for s in range(0, count, batch):
            chunk = params[s:s + batch]
            losses = np.asarray(evaluate_params(chunk), dtype=np.float64)
            for p, loss in zip(chunk, losses):
                history.append({"params": {k: float(p[k]) for k, _, _ in free}, "loss": float(loss)})
                full_params.append(p)

    n_local = int(round(n_trials * refine_fraction)) if free else 0
    run([(lo, hi) for _, lo, hi in free], n_trials - n_local, seed)
    best = min(range(len(history)), key=lambda i: (history[i]["loss"], i))
    if n_local:
        centre = full_params[best]
        box = []
        for key, lo, hi in free:
            half = 0.1 * (hi - lo)
            box.append((max(lo, centre[key] - half), min(hi, centre[key] + half)))
        run(box, n_local, seed + 1)
        best = min(range(len(history)), key=lambda i: (history[i]["loss"], i))
    best_params = {k: float(full_params[best][k]) for k, _, _ in free}
    return best_params, float(history[best]["loss"]), history


def _quadratic_terms(u: np.ndarray) -> np.ndarray:
    """Design matrix [1, u_i, u_i u_j (i <= j)] of a full quadratic in d variables."""
    k, d = u.shape
    cols = [np.ones(k)] + [u[:, i] for i in range(d)]
    cols += [u[:, i] * u[:, j] for i in range(d) for j in range(i, d)]
    return np.stack(cols, axis=1)


def _quadratic_minimiser(u: np.ndarray, f: np.ndarray):
    """Least-squares quadratic through (u, f); returns the Newton step of the
    model with its Hessian's eigenvalues replaced by their magnitudes (floored
    at 1e-3 of the largest), so a saddle or a ridge still gives a descent
    step, or None when the fit is degenerate."""
    d = u.shape[1]
    A = _quadratic_terms(u)
    if np.linalg.matrix_rank(A) < A.shape[1]:
        return None
    coef, *_ = np.linalg.lstsq(A, f, rcond=None)
    g = coef[1:1 + d]
    H = np.empty((d, d))
    idx = 1 + d
    for i in range(d):
        for j in range(i, d):
            H[i, j] = H[j, i] = coef[idx] * (2.0 if i == j else 1.0)
            idx += 1
    w, V = np.linalg.eigh(H)
    top = float(np.max(np.abs(w)))
    if not np.isfinite(top) or top == 0.0:
        return None
    w = np.maximum(np.abs(w), 1e-3 * top)
    return -(V @ ((V.T @ g) / w))


def quadratic_refine(evaluate: Evaluate, x0, radius, rounds: int = 3, batch: int = 16, shrink: float = 0.3,
                     seed: int = 0, lo=None, hi=None) -> tuple[np.ndarray, float, int]:
    """Batched trust-region refinement from ``x0``; returns (best_x, best_loss, n_evals).

    Round r evaluates ``batch`` points of the box centre +/- radius_r in one
    ``evaluate`` call: the centre, the 2d axis points and a scrambled Sobol
    fill.  A full quadratic (1 + d + d(d+1)/2 coefficients) is fitted to
    every point seen so far inside the box, in the box's unit coordinates;
    its Newton step (Hessian eigenvalues taken by magnitude, so an indefinite
    model still yields a descent step), pulled back into the box, is
    evaluated on its own.  The next centre is the best point seen so far.
    When the model's point is that best point (the model is trusted) the box
    scales by max(shrink, 2 x the step's reach in box units), at most 1;
    otherwise it halves.  ``lo``/``hi`` clip every candidate.  Ties keep the
    earliest point.
    """
    from scipy.stats import qmc

    x0 = np.asarray(x0, dtype=np.float64).reshape(-1)
    radius = np.broadcast_to(np.asarray(radius, dtype=np.float64), x0.shape).copy()
    d = x0.size
    need = 1 + d + d * (d + 1) // 2
    if batch < max(need, 1 + 2 * d):
        raise ValueError(f"batch must hold at least {max(need, 1 + 2 * d)} points for {d} parameters")
    if np.any(radius <= 0):
        raise ValueError("radius must be positive")
    lo = np.full(d, -np.inf) if lo is None else np.asarray(lo, dtype=np.float64)
    hi = np.full(d, np.inf) if hi is None else np.asarray(hi, dtype=np.float64)
    sobol = qmc.Sobol(d=d, scramble=True, seed=seed)
    fill = 2.0 * sobol.random_base2(int(math.ceil(math.log2(max(batch, 2))))) - 1.0
    axes = np.concatenate([np.eye(d), -np.eye(d)])
    unit = np.concatenate([np.zeros((1, d)), axes, fill[:batch - 1 - 2 * d]])
    xs: list[np.ndarray] = []
    fs: list[float] = []
    centre = np.clip(x0, lo, hi)
    for _ in range(rounds):
        pts = np.clip(centre + unit * radius, lo, hi)
        f = np.asarray(evaluate(pts), dtype=np.float64).reshape(-1)
        xs.extend(pts)
        fs.extend(float(v) for v in f)
        # every point seen so far inside the current box enters the fit
        u_all = (np.asarray(xs) - centre) / radius
        inside = np.max(np.abs(u_all), axis=1) <= 1.0 + 1e-12
        step = _quadratic_minimiser(u_all[inside], np.asarray(fs)[inside])
        scale = 0.5
        if step is not None and np.all(np.isfinite(step)):
            reach = float(np.max(np.abs(step)))
            step /= max(1.0, reach)
            xm = np.clip(centre + step * radius, lo, hi)
            fm = float(np.asarray(evaluate(xm[None, :]), dtype=np.float64).reshape(-1)[0])
            xs.append(xm)
            fs.append(fm)
            if fm <= min(fs):
                scale = min(1.0, max(shrink, 2.0 * min(reach, 1.0)))
        centre = xs[int(np.argmin(fs))].copy()
        radius = radius * scale
    best = int(np.argmin(fs))
    return np.asarray(xs[best]), float(fs[best]), len(fs)
