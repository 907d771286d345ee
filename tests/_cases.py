"""Seeded inputs shared by the oracle tests and the GPU parity tests.

Each builder re-creates an input that a reference test constructs (cited),
or a synthetic benchmark dataset, on the CPU with NumPy.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from oracle import ssb_oracle as O

MAKE_ENGINE_ABERRATIONS = (-120.0, 55.0, math.radians(17.0))


@dataclass
class EngineCase:
    """Inputs of the reference's ``_make_engine`` (``tests/test_ssb_cuda_128.py:212-249``)."""
    G: np.ndarray               # complex64 [b, N, N] or Hermitian [b, N, N/2+1]
    rows: np.ndarray
    cols: np.ndarray
    center: tuple[float, float]
    dc: complex
    n: int
    det: tuple[int, int]
    recip: tuple[float, float] | None      # None: ang/wavelength, as a backend derives it
    wavelength: float
    semiangle_mrad: float
    ang_mrad: tuple[float, float]
    q_step: float

    def geometry(self, rotation_rad: float = 0.0, dtype=np.float64) -> O.Geometry:
        return O.detector_geometry(self.rows, self.cols, self.center, self.det, self.ang_mrad,
                                   self.wavelength, self.semiangle_mrad, rotation_rad, self.n,
                                   self.q_step, dtype, recip=self.recip)


def make_engine_case(size: int = 128, num_bf: int = 7, g_qk=None,
                     bf_center=(15.5, 15.5)) -> EngineCase:
    if g_qk is None:
        rng = np.random.default_rng(1234)
        real = rng.standard_normal((num_bf, size, size), dtype=np.float32)
        imag = rng.standard_normal((num_bf, size, size), dtype=np.float32)
        g_qk = (real + 1j * imag).astype(np.complex64)
    g_qk = np.asarray(g_qk, dtype=np.complex64)
    row_pattern = np.asarray([13, 14, 15, 16, 17, 16, 15], dtype=np.int32)
    col_pattern = np.asarray([14, 15, 16, 17, 16, 15, 14], dtype=np.int32)
    rows = np.resize(row_pattern, num_bf).astype(np.int32)
    cols = np.resize(col_pattern, num_bf).astype(np.int32)
    dc = complex(g_qk[:, 0, 0].astype(np.complex128).mean())
    return EngineCase(g_qk, rows, cols, (float(bf_center[0]), float(bf_center[1])), dc, size,
                      (32, 32), (1.0 / 32.0, 1.0 / 32.0), 0.0197, 21.4, (1.0, 1.0), 0.5)


def hermitian_stack(size: int, num_bf: int) -> np.ndarray:
    """``fft2`` of a seeded real stack (``tests/test_ssb_cuda_128.py:481-491``), full plane."""
    rng = np.random.default_rng(4321)
    real = rng.standard_normal((num_bf, size, size), dtype=np.float32)
    return np.fft.fft2(real.astype(np.float64)).astype(np.complex64)


def poisson_data(seed: int, scan=(128, 128), det=16, radius=4, add=80) -> np.ndarray:
    """``poisson(4, (scan, scan, 16, 16))`` plus ``add`` inside a disk of radius 4
    at (8, 8) (``tests/test_ssb_cuda_128.py:591-595``, ``654-658``, ``951-955``)."""
    rng = np.random.default_rng(seed)
    data = rng.poisson(4.0, size=(scan[0], scan[1], det, det)).astype(np.uint16)
    yy, xx = np.ogrid[:det, :det]
    bf = (yy - det // 2) ** 2 + (xx - det // 2) ** 2 <= radius ** 2
    data[..., bf] += add
    return data


POISSON_KW = dict(voltage_kV=300, semiangle_mrad=21.4, scan_sampling=0.5, det_sampling=1.0)


def ramp_data(n: int = 128, det: int = 8) -> np.ndarray:
    """``I[r, c, i, j] = r + 2c + 8i + j`` as uint16 (closed-form DFT)."""
    r = np.arange(n, dtype=np.int64)[:, None, None, None]
    c = np.arange(n, dtype=np.int64)[None, :, None, None]
    i = np.arange(det, dtype=np.int64)[None, None, :, None]
    j = np.arange(det, dtype=np.int64)[None, None, None, :]
    return (r + 2 * c + 8 * i + j).astype(np.uint16)


def phase_errors(phase, ref) -> tuple[float, float, float]:
    """(RMS, p99.9, max) of the wrapped phase difference in radians."""
    d = np.angle(np.exp(1j * (np.asarray(phase, np.float64) - np.asarray(ref, np.float64))))
    a = np.abs(d)
    return float(np.sqrt(np.mean(d * d))), float(np.percentile(a, 99.9)), float(a.max())
