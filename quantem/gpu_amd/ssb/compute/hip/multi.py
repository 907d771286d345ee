"""Batched aberration search over several datasets at once, including
datasets of uneven scan size (each scan is padded or cropped to its
supported size by its own backend, as the reference does,
``cuda/backend.py:229-269``).

Every candidate of one seeded Sobol design is evaluated on every dataset:
for each batch, all datasets' exact-loss launches are enqueued back to back
on the stream and the host synchronises once, so the datasets share one
launch sequence per batch instead of one search each.  Returns each
dataset's best candidate; each can then be refined on its own.  No
reference counterpart (the reference searches one dataset per call,
``cuda/optimizer.py:59-230``).
"""
from __future__ import annotations

import math

import numpy as np

from . import optimizer as _opt


def evaluate_many(backends, params: list[dict]) -> np.ndarray:
    """[datasets, candidates] exact losses, one synchronisation."""
    import torch
    pending = [b._enqueue_params(params) for b in backends]
    torch.cuda.synchronize()
    return np.stack([p.cpu().numpy().astype(np.float64) for p in pending])


def search_datasets(backends, aberrations: dict, n_trials: int = 512, seed: int = 42, batch: int = 64,
                    keys=_opt.PARAM_KEYS + ("C30", "rotation_angle_deg")) -> list[dict]:
    """Per dataset: {"params", "loss", "history"} of the shared design's best
    candidate (ties keep the earliest)."""
    from scipy.stats import qmc

    if not backends:
        return []
    free, fixed = _opt._split_specs(dict(aberrations), {k: 0.0 for k in keys}, keys)
    if free:
        unit = qmc.Sobol(d=len(free), scramble=True, seed=seed).random_base2(
            int(math.ceil(math.log2(max(n_trials, 1)))))[:n_trials]
    else:
        unit = np.zeros((n_trials, 0))
    params = []
    for u in unit:
        p = dict(fixed)
        for (key, lo, hi), v in zip(free, u):
            p[key] = lo + (hi - lo) * float(v)
        params.append(p)
    losses = np.concatenate([evaluate_many(backends, params[s:s + batch]) for s in range(0, n_trials, batch)],
                            axis=1)
    out = []
    for d in range(len(backends)):
        best = int(np.argmin(losses[d]))
        out.append({"params": dict(params[best]), "loss": float(losses[d, best]),
                    "history": [{"params": {k: float(p[k]) for k, _, _ in free}, "loss": float(l)}
                                for p, l in zip(params, losses[d])]})
    return out
