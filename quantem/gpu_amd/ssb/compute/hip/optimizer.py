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
* ``quadratic_refine`` is a batched local refinement: each round
  evaluates one batch (the centre, the axis points and a Sobol fill of a
  box), fits a full quadratic model by least squares and moves to its
  minimiser inside the box; it needs a few batches where Nelder-Mead needs
  100+ sequential evaluations.  It has no reference counterpart.
* ``newton_refine`` is the batched finite-difference Newton refinement
  (the fits' default): one 10-point stencil per iteration, a trust-region
  step of the quadratic it interpolates, a stencil floor above the loss's
  noise.  It has no reference counterpart.
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


def model_polish(evaluate: Evaluate, x0, box, rounds: int = 2, batch: int = 16, seed: int = 20, clip: float = 1.5,
                 lo=None, hi=None) -> tuple[np.ndarray, int]:
    """Least-squares quadratic polish near an optimum; returns (x, n_evals).

    Round r evaluates ``batch`` scrambled-Sobol points of the box
    ``x +/- box`` in one ``evaluate`` call, fits a full quadratic to them in
    box units and moves ``x`` to the model's minimiser (the Newton step with
    the Hessian's eigenvalues taken by magnitude, at most ``clip`` boxes per
    axis).  The result is the last model minimiser, not the lowest loss seen:
    where the objective is rough at the scale of the answer (the exact loss
    moves in small steps as overlap points cross the aperture edge), a fit
    through many points averages the roughness that picking the lowest
    sample would select.  The minimiser itself is not evaluated.  ``lo``/``hi``
    clip every point.  It has no reference counterpart.
    """
    from scipy.stats import qmc

    x = np.asarray(x0, dtype=np.float64).reshape(-1).copy()
    d = x.size
    box = np.broadcast_to(np.asarray(box, dtype=np.float64), x.shape).copy()
    need = 1 + d + d * (d + 1) // 2
    if batch < need:
        raise ValueError(f"batch must hold at least {need} points for {d} parameters")
    if np.any(box <= 0):
        raise ValueError("box must be positive")
    lo = np.full(d, -np.inf) if lo is None else np.asarray(lo, dtype=np.float64)
    hi = np.full(d, np.inf) if hi is None else np.asarray(hi, dtype=np.float64)
    x = np.clip(x, lo, hi)
    n = 0
    for r in range(rounds):
        u = 2.0 * qmc.Sobol(d=d, scramble=True, seed=seed + r).random_base2(int(math.ceil(math.log2(batch))))[:batch] - 1.0
        pts = np.clip(x + u * box, lo, hi)
        f = np.asarray(evaluate(pts), dtype=np.float64).reshape(-1)
        n += len(f)
        step = _quadratic_minimiser((pts - x) / box, f)
        if step is None or not np.all(np.isfinite(step)):
            continue
        x = np.clip(x + np.clip(step, -clip, clip) * box, lo, hi)
    return x, n


def _stencil(d: int) -> np.ndarray:
    """Unit offsets of the finite-difference stencil: the centre, +/- each
    axis, and +i+j for every pair (1 + 2d + d(d-1)/2 points)."""
    rows = [np.zeros(d)]
    eye = np.eye(d)
    rows += [eye[i] for i in range(d)] + [-eye[i] for i in range(d)]
    rows += [eye[i] + eye[j] for i in range(d) for j in range(i + 1, d)]
    return np.asarray(rows)


def _stencil_model(f: np.ndarray, h: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Gradient and Hessian at the centre from the stencil losses ``f``
    (order of ``_stencil``) with axis steps ``h``: central first and second
    differences on the axes, (f_ij - f_i - f_j + f_0) / (h_i h_j) across."""
    d = h.size
    f0, fp, fm = f[0], f[1:1 + d], f[1 + d:1 + 2 * d]
    g = (fp - fm) / (2.0 * h)
    H = np.diag((fp - 2.0 * f0 + fm) / (h * h))
    k = 1 + 2 * d
    for i in range(d):
        for j in range(i + 1, d):
            H[i, j] = H[j, i] = (f[k] - fp[i] - fp[j] + f0) / (h[i] * h[j])
            k += 1
    return g, H


def newton_refine(evaluate: Evaluate, x0, step, max_iter: int = 8, xtol=None, hmin=None, lo=None, hi=None,
                  trust: float = 4.0, ftol: float = 0.0, polish: bool = False) -> tuple[np.ndarray, float, int]:
    """Batched finite-difference Newton refinement from ``x0``; returns
    (best_x, best_loss, n_evals).

    Each iteration evaluates one batch: the stencil of ``_stencil`` around
    the centre with axis steps ``h = max(tau * step, hmin)`` (``tau`` starts
    at 1; ``hmin``, default ``step / 20``, keeps the differences above the
    loss's own noise).  The model it interpolates, in stencil units
    ``u = (x - centre) / h`` (so the parameters' scales do not enter its
    conditioning), gives the Newton step -- Hessian eigenvalues taken by
    magnitude and floored at 1e-6 of the largest, so a saddle or a nearly
    flat valley still gives a bounded descent step -- scaled down, keeping
    its direction, to end inside the box ``|u_i| <= trust``.  The step's end
    is the next centre; the next batch both checks it and models around it.
    With rho = actual / predicted decrease: an uphill step (the centre worse
    than the best point seen) restarts from the best point with ``tau / 4``;
    rho < 1/4 halves ``tau``; rho > 3/4 on a step that reached the box edge
    doubles it (at most 4); a step inside half the box shrinks ``tau``
    towards ``2 |s| / step`` so the stencil follows the distance still to
    go.  Stops when a step has every ``|s_i| < xtol_i`` (default
    ``step / 100``), when its predicted decrease is below ``ftol`` (the
    loss's noise: no later batch could confirm it), or after ``max_iter``
    batches.  ``polish`` then fits a least-squares quadratic to every point
    within two stencils of the best and evaluates its minimiser (one more
    point).  The result is the best point evaluated.  Every point stays in
    ``[lo, hi]``: the stencil keeps its symmetric shape (half-width at most
    half the box) and its centre moves inward when it would cross a bound,
    so the differences stay central; steps are clipped to the box.  It has
    no reference counterpart.
    """
    x0 = np.asarray(x0, dtype=np.float64).reshape(-1)
    d = x0.size
    step = np.broadcast_to(np.asarray(step, dtype=np.float64), x0.shape).copy()
    if np.any(step <= 0):
        raise ValueError("step must be positive")
    xtol = step / 100.0 if xtol is None else np.broadcast_to(np.asarray(xtol, dtype=np.float64), x0.shape).copy()
    hmin = step / 20.0 if hmin is None else np.broadcast_to(np.asarray(hmin, dtype=np.float64), x0.shape).copy()
    lo = np.full(d, -np.inf) if lo is None else np.asarray(lo, dtype=np.float64)
    hi = np.full(d, np.inf) if hi is None else np.asarray(hi, dtype=np.float64)
    unit = _stencil(d)
    xs: list[np.ndarray] = []
    fs: list[float] = []
    target = np.clip(x0, lo, hi)   # the point the next batch checks (the last step's end)
    tau = 1.0
    last_f = math.inf           # loss at the point the last step started from
    pred = 0.0                  # the model's predicted decrease of that step
    edge = False                # that step reached the trust box
    n_unit = len(unit)
    for _ in range(max_iter):
        # a symmetric stencil that fits the box: at most half its width per
        # axis, its centre moved inward when target +- h would cross a bound
        # (a stencil clipped at a bound would collapse one side onto the
        # centre); the target itself is then evaluated as one extra point
        h = np.minimum(np.maximum(tau * step, hmin), (hi - lo) / 2.0)
        if np.any(h <= 0):
            break
        centre = np.clip(target, lo + h, hi - h)
        pts = centre + unit * h
        shifted = not np.array_equal(centre, target)
        if shifted:
            pts = np.vstack([pts, target[None, :]])
        f = np.asarray(evaluate(pts), dtype=np.float64).reshape(-1)
        xs.extend(pts)
        fs.extend(float(v) for v in f)
        f_target = float(f[-1] if shifted else f[0])
        if f_target > last_f:
            # the last step went uphill: model again around the best point
            target = xs[int(np.argmin(fs))].copy()
            last_f = min(fs)
            tau /= 4.0
            pred = 0.0
            continue
        if pred > 0.0:
            rho = (last_f - f_target) / pred
            if rho < 0.25:
                tau /= 2.0
            elif rho > 0.75 and edge:
                tau = min(2.0 * tau, 4.0)
        last_f = f_target
        g, H = _stencil_model(f[:n_unit], h)
        gu, Hu = g * h, H * np.outer(h, h)                               # stencil units
        w, V = np.linalg.eigh(Hu)
        top = float(np.max(np.abs(w)))
        if not np.isfinite(top) or top == 0.0 or not np.all(np.isfinite(gu)):
            break
        w = np.maximum(np.abs(w), 1e-6 * top)
        gv = V.T @ gu
        su = -(V @ (gv / w))
        reach = float(np.max(np.abs(su))) / trust
        edge = reach >= 1.0
        if edge:
            su = su / reach          # keep the direction, end on the box
        new_target = np.clip(centre + su * h, lo, hi)
        moved = new_target - target
        if np.all(np.abs(moved) < xtol):
            break

        def model(u):
            v = V.T @ u
            return float(gv @ v + 0.5 * np.sum(w * v * v))
        pred = model((target - centre) / h) - model((new_target - centre) / h)
        if pred < ftol:
            break
        if not edge and reach < 0.5:
            tau = min(tau, 2.0 * float(np.max(np.abs(moved) / step)))
        target = new_target
    if polish and fs:
        # least-squares quadratic through every point within 2 stencils of the
        # best one: many points average the loss's noise that the last
        # stencil's interpolation takes in full; its minimiser costs one more
        # evaluation and is kept only if it is better
        b0 = int(np.argmin(fs))
        h = np.maximum(tau * step, hmin)
        u = (np.asarray(xs) - xs[b0]) / h
        near = np.max(np.abs(u), axis=1) <= 2.0
        su = _quadratic_minimiser(u[near], np.asarray(fs)[near]) if np.count_nonzero(near) >= 2 * len(unit) else None
        if su is not None and np.all(np.isfinite(su)):
            su = su / max(1.0, float(np.max(np.abs(su))) / 2.0)
            xm = np.clip(xs[b0] + su * h, lo, hi)
            fm = float(np.asarray(evaluate(xm[None, :]), dtype=np.float64).reshape(-1)[0])
            xs.append(xm)
            fs.append(fm)
    best = int(np.argmin(fs))
    return np.asarray(xs[best]), float(fs[best]), len(fs)
