"""Batched aberration search with the phase-image-variance objective.

``TrialBatch`` wraps the native ``ssb_trials_*`` handle: T trial aberration
sets (C10, C12/phi12, C30 and the other even-order terms, each with its own
scan rotation) are evaluated on one or more datasets -- engines holding a
resident Hermitian Fourier stack, scans of different size allowed -- in one
persistent launch, optionally replayed from a captured hipGraph.  The
objective is ``var_R angle(O_t(R))`` of each trial's Fourier-sum object
(``cuda/engine.py:1047-1107`` per trial), a contrast measure that the
correct aberrations maximise.  It replaces the reference's per-trial loop
(``cuda/engine.py:1813-1836``, ``cuda/optimizer.py:59-191``) for that
objective; the reference's default exact objective stays on
``HipSSBEngine.batch_phase_loss``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi


class TrialBatch:
    """A batched phase-variance evaluator over ``engines`` (each after
    ``prepare_g``/``import_g`` with a Hermitian stack) for up to
    ``max_trials`` trials per call."""

    def __init__(self, engines, max_trials: int):
        import torch
        self.torch = torch
        self.engines = list(engines)
        if not self.engines:
            raise ValueError("at least one engine is needed")
        dev = self.engines[0].plan.device
        if any(e.plan.device != dev for e in self.engines):
            raise ValueError("all engines must live on one device")
        self.device = dev
        self.max_trials = int(max_trials)
        handles = (ctypes.c_void_p * len(self.engines))(*[e.plan.handle.value for e in self.engines])
        out = ctypes.c_void_p()
        with torch.cuda.device(dev):
            _abi.call("ssb_trials_create", ctypes.cast(handles, _abi._pp), len(self.engines), self.max_trials,
                      ctypes.byref(out))
        self._h = out
        self.n_trials = 0
        self.out = torch.empty((len(self.engines), self.max_trials), dtype=torch.float32, device=dev)

    def _stream(self):
        return _abi.stream_handle(self.torch.cuda.current_stream(self.device))

    def set(self, mags, angles, rotations=None) -> None:
        """Trial aberration vectors ``[T, 14]`` (layout of ``batch_phase_loss``)
        and scan rotations ``[T]`` in radians."""
        m = np.ascontiguousarray(np.asarray(mags, dtype=np.float64).reshape(-1, 14))
        a = np.ascontiguousarray(np.asarray(angles, dtype=np.float64).reshape(-1, 14))
        if m.shape != a.shape:
            raise ValueError("mags and angles must have the same shape")
        rot = None
        if rotations is not None:
            rot = np.ascontiguousarray(np.asarray(rotations, dtype=np.float64).reshape(-1))
            if rot.size != m.shape[0]:
                raise ValueError("one rotation per trial is needed")
        with self.torch.cuda.device(self.device):
            _abi.call("ssb_trials_set", self._h, int(m.shape[0]), m.ctypes.data_as(_abi._dp),
                      a.ctypes.data_as(_abi._dp), rot.ctypes.data_as(_abi._dp) if rot is not None else None,
                      self._stream())
        self.n_trials = int(m.shape[0])

    def phase_variance(self, use_graph: bool = True):
        """``[datasets, T]`` phase-image variances (device tensor, asynchronous)."""
        with self.torch.cuda.device(self.device):
            _abi.call("ssb_trials_phase_variance", self._h, _abi.ptr(self.out), 1 if use_graph else 0,
                      self._stream())
        return self.out[:, :self.n_trials]

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _abi.lib().ssb_trials_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def sobol_design(aberrations: dict, n_trials: int, seed: int = 42,
                 keys=("C10_nm", "C12_nm", "phi12_deg", "C30", "rotation_angle_deg")) -> list[dict]:
    """``n_trials`` candidates of a scrambled Sobol design over the ranges in
    ``aberrations`` (``(lo, hi)`` pairs; scalars stay fixed, missing keys are 0)."""
    from scipy.stats import qmc
    free = [(k, float(v[0]), float(v[1])) for k, v in aberrations.items() if isinstance(v, (tuple, list))]
    fixed = {k: 0.0 for k in keys}
    fixed.update({k: float(v) for k, v in aberrations.items() if not isinstance(v, (tuple, list))})
    unit = (qmc.Sobol(d=len(free), scramble=True, seed=seed).random_base2(int(math.ceil(math.log2(max(n_trials, 1)))))
            [:n_trials] if free else np.zeros((n_trials, 0)))
    out = []
    for u in unit:
        p = dict(fixed)
        for (key, lo, hi), v in zip(free, u):
            p[key] = lo + (hi - lo) * float(v)
        out.append(p)
    return out


def params_to_vectors(params: list[dict]):
    """(mags [T, 14], angles [T, 14], rotations [T] rad) of search candidates,
    rounded to float32 as the exact objective's candidates are."""
    n = len(params)
    mags, angs, rot = np.zeros((n, 14)), np.zeros((n, 14)), np.zeros(n)
    for i, p in enumerate(params):
        mags[i, 0] = np.float32(p.get("C10_nm", 0.0))
        mags[i, 1] = np.float32(p.get("C12_nm", 0.0))
        angs[i, 1] = np.float32(math.radians(p.get("phi12_deg", 0.0)))
        mags[i, 4] = np.float32(p.get("C30", 0.0))
        rot[i] = math.radians(p.get("rotation_angle_deg", 0.0))
    return mags, angs, rot


def search_phase_variance(engines, aberrations: dict, n_trials: int = 512, seed: int = 42,
                          use_graph: bool = True, batch: TrialBatch | None = None) -> list[dict]:
    """One batched launch of a shared Sobol design over every dataset; per
    dataset the candidate of largest phase-image variance (ties keep the
    earliest): {"params", "phase_variance", "history"}."""
    params = sobol_design(aberrations, n_trials, seed)
    own = batch is None
    batch = batch or TrialBatch(engines, n_trials)
    try:
        batch.set(*params_to_vectors(params))
        var = batch.phase_variance(use_graph).cpu().numpy().astype(np.float64)
    finally:
        if own:
            batch.close()
    out = []
    for d in range(var.shape[0]):
        best = int(np.argmax(var[d]))
        out.append({"params": dict(params[best]), "phase_variance": float(var[d, best]),
                    "history": [{"params": dict(p), "phase_variance": float(v)} for p, v in zip(params, var[d])]})
    return out
