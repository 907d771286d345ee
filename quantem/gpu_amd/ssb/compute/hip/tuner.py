"""Autotuner of the reconstruction path's launch shapes.

Three knobs of the native library (``ssb_tune_*`` in ``include/quantem_ssb.h``)
change launch shapes only -- the workgroup count of the raw-data stream pass
(exact integer sums, bit-identical), the number of FFT pairs per gather
workgroup and the tile shape of the fused column FFT + trotter (the same
arithmetic in another instantiation, equal to float rounding).  ``autotune`` times every
candidate of each knob in turn (coordinate descent, the others held at their
best so far) on the caller's own resident data with the per-stage hipEvents
of ``ssb_recon_u16``, keeps the fastest, and can store the choice in a JSON
cache keyed by device, scan size, detector size and BF count so later runs
reuse it (``load_cached``).  The reference has no counterpart: its launch
shapes are fixed per scan size (``cuda/kernels/fft256.py:1051-1058``).
"""
from __future__ import annotations

import ctypes
import json
import os
import time
from pathlib import Path

import numpy as np

from . import _abi

KNOBS = {"stream_wg": 0, "gather_pairs": 1, "trotter": 2, "split": 4}
# the recon stages whose time each knob changes
STAGE = {"stream_wg": ("mean_dp_stream",), "gather_pairs": ("gather_rowfft",), "trotter": ("colfft_trotter",),
         "split": ("gather_rowfft", "colfft_trotter")}


def candidates(scan_n: int, knob: str) -> list[int]:
    """Allowed values of a knob at this scan size; the first is the default."""
    vals = (ctypes.c_int * 8)()
    n = _abi.lib().ssb_tune_candidates(int(scan_n), KNOBS[knob], vals, 8)
    if n < 0:
        _abi.check(n, "ssb_tune_candidates")
    return [int(vals[i]) for i in range(n)]


def cache_path() -> Path:
    env = os.environ.get("QSSB_TUNE_CACHE")
    if env:
        return Path(env)
    return Path.home() / ".cache" / "quantem_gpu_amd" / "tuning.json"


def cache_key(device_name: str, scan_n: int, det: tuple[int, int], nbf: int) -> str:
    """BF counts share an entry per power-of-two bucket."""
    bucket = 1 << max(0, int(nbf) - 1).bit_length()
    return f"{device_name}|scan{scan_n}|det{det[0]}x{det[1]}|bf{bucket}"


def _read_cache(path: Path) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def load_cached(engine, nbf: int, path: Path | None = None) -> dict | None:
    """Apply a cached choice for this device/shape, if one exists; returns it."""
    torch = engine.torch
    s = engine.plan.spec
    key = cache_key(torch.cuda.get_device_name(engine.plan.device), s.scan_n, (s.det_rows, s.det_cols), nbf)
    entry = _read_cache(path or cache_path()).get(key)
    if not entry:
        return None
    for knob, value in entry.items():
        if knob in KNOBS:
            engine.tune_set(KNOBS[knob], int(value))
    return entry


def _time_stage(engine, run, stages, repeats: int) -> float:
    """Median per-call milliseconds of the recon stages ``stages`` (summed)
    over ``repeats`` calls."""
    run()                                   # warm the new launch shape
    times = []
    for _ in range(repeats):
        engine.timing_begin(1)
        run()
        t = engine.timing_end()
        times.append(sum(t[st] for st in stages))
    return float(np.median(times))


def coordinate_descent(measure, knob_values: dict, defaults: dict) -> tuple[dict, dict]:
    """Pick each knob's fastest value in turn, the others at their best so far.

    ``measure(knob, value, config)`` returns a time; ties keep the earlier
    (default-first) candidate.  Returns (best config, all measurements)."""
    best = dict(defaults)
    table: dict = {}
    for knob, values in knob_values.items():
        timings = []
        for v in values:
            cfg = dict(best)
            cfg[knob] = v
            timings.append(measure(knob, v, cfg))
        table[knob] = dict(zip(values, timings))
        best[knob] = values[int(np.argmin(timings))]
    return best, table


def autotune(engine, data, c10: float = 0.0, repeats: int = 5, knobs=tuple(KNOBS), store: bool = False,
             path: Path | None = None) -> dict:
    """Tune ``engine``'s knobs on resident raw ``data`` (uint16, GPU); returns
    {"config": {knob: value}, "ms": {knob: {value: stage ms}}, "seconds": t}.
    The chosen values stay set on the engine."""
    t0 = time.perf_counter()
    n = engine.plan.spec.scan_n
    knob_values = {k: candidates(n, k) for k in knobs}
    defaults = {k: v[0] for k, v in knob_values.items()}

    def run():
        engine.recon_u16(data, c10, 0.0, 0.0)

    def measure(knob, value, cfg):
        for k, v in cfg.items():
            engine.tune_set(KNOBS[k], v)
        return _time_stage(engine, run, STAGE[knob], repeats)

    best, table = coordinate_descent(measure, knob_values, defaults)
    for k, v in best.items():
        engine.tune_set(KNOBS[k], v)
    if store:
        torch = engine.torch
        s = engine.plan.spec
        key = cache_key(torch.cuda.get_device_name(engine.plan.device), n, (s.det_rows, s.det_cols), engine.num_bf)
        p = path or cache_path()
        entries = _read_cache(p)
        entries[key] = best
        p.parent.mkdir(parents=True, exist_ok=True)
        p.write_text(json.dumps(entries, indent=1, sort_keys=True))
    return {"config": best, "ms": {k: {str(v): t for v, t in d.items()} for k, d in table.items()},
            "seconds": time.perf_counter() - t0}
