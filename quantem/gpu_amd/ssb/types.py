"""Backend-neutral SSB types, mirroring the reference's contract so a ``hip``
backend is interchangeable with the CUDA one.

* ``SSBPrecision``, ``SSBExportState``, ``SSBProtocol`` --
  ``src/quantem/gpu/ssb/compute/protocol.py:15-125``;
* ``BrightfieldDisk`` -- ``src/quantem/gpu/ssb/bf_selector.py:11-71``;
* ``SSBResult`` -- ``src/quantem/gpu/ssb/results.py:41-129``.

Field names, meanings and validation messages follow the reference; arrays
may be torch tensors on the GPU (``phase``/``amplitude`` then stay there).
"""
from __future__ import annotations

import math
import pathlib
from contextlib import AbstractContextManager
from dataclasses import dataclass, field
from typing import Literal, Protocol, runtime_checkable

import numpy as np

Backend = Literal["cuda", "mps", "webgpu", "hip"]


@dataclass(frozen=True)
class SSBPrecision:
    """Numeric storage of every SSB backend (``protocol.py:15-20``)."""
    real_dtype: Literal["float32"] = "float32"
    complex_dtype: Literal["complex64"] = "complex64"


@dataclass(frozen=True)
class BrightfieldDisk:
    """Validated, read-only BF detector coordinates (``bf_selector.py:11-71``)."""
    rows: np.ndarray
    cols: np.ndarray
    center_row_col: tuple[float, float]
    radius_px: float
    detected_radius_px: float
    detector_shape: tuple[int, int]

    def __post_init__(self) -> None:
        rows = np.asarray(self.rows, dtype=np.int32).reshape(-1).copy()
        cols = np.asarray(self.cols, dtype=np.int32).reshape(-1).copy()
        if rows.size == 0 or rows.shape != cols.shape:
            raise ValueError("BF rows and columns must be non-empty matching vectors.")
        shape = tuple(int(v) for v in self.detector_shape)
        if len(shape) != 2 or min(shape) < 1:
            raise ValueError("detector_shape must contain two positive (row, col) sizes.")
        if (int(rows.min()) < 0 or int(cols.min()) < 0 or int(rows.max()) >= shape[0]
                or int(cols.max()) >= shape[1]):
            raise ValueError(f"BF coordinates fall outside detector shape {shape}.")
        if np.unique(rows.astype(np.int64) * shape[1] + cols).size != rows.size:
            raise ValueError("BF coordinates must not contain duplicates.")
        center = tuple(float(v) for v in self.center_row_col)
        if len(center) != 2 or not all(math.isfinite(v) for v in center):
            raise ValueError("center_row_col must contain two finite values.")
        radius = float(self.radius_px)
        detected = float(self.detected_radius_px)
        if not math.isfinite(radius) or radius <= 0:
            raise ValueError("radius_px must be a positive finite value.")
        if not math.isfinite(detected) or detected <= 0:
            raise ValueError("detected_radius_px must be a positive finite value.")
        rows.flags.writeable = False
        cols.flags.writeable = False
        object.__setattr__(self, "rows", rows)
        object.__setattr__(self, "cols", cols)
        object.__setattr__(self, "center_row_col", center)
        object.__setattr__(self, "radius_px", radius)
        object.__setattr__(self, "detected_radius_px", detected)
        object.__setattr__(self, "detector_shape", shape)

    @property
    def size(self) -> int:
        return int(self.rows.size)


@dataclass(frozen=True)
class SSBExportState:
    """Compact host state for browser/widget consumers (``protocol.py:23-58``)."""
    backend: Backend
    scan_shape: tuple[int, int]
    brightfield: BrightfieldDisk
    kx_bf: np.ndarray
    ky_bf: np.ndarray
    qx_1d: np.ndarray
    qy_1d: np.ndarray
    aperture_k: np.ndarray
    alpha_k2: np.ndarray
    cos2phi_k: np.ndarray
    sin2phi_k: np.ndarray
    wavelength_A: float
    semiangle_rad: float
    angular_sampling_rad: tuple[float, float]
    sampling_A: tuple[float, float]
    dc_value: complex
    precision: SSBPrecision = SSBPrecision()
    bf_source_path: pathlib.Path | None = None
    bf_source_dtype: np.dtype | None = None
    bf_source_max_value: int | None = None

    @property
    def num_bf(self) -> int:
        return self.brightfield.size

    @property
    def active_num_bf(self) -> int:
        return int(np.count_nonzero(np.asarray(self.aperture_k) > 0.0))


def _is_torch(x) -> bool:
    try:
        import torch
    except ModuleNotFoundError:
        return False
    return isinstance(x, torch.Tensor)


@dataclass
class SSBResult:
    """Result of one SSB reconstruction (``results.py:41-129``)."""
    object_wave: object
    backend: Backend
    aberrations: dict[str, float] = field(default_factory=dict)
    rotation_angle_deg: float = 0.0
    loss: float | None = None
    elapsed: float | None = None
    timings: dict[str, float] = field(default_factory=dict)
    n_trials: int | None = None
    num_bf: int | None = None
    refine_method: str | None = None
    refine_nfev: int | None = None
    refine_elapsed: float | None = None
    voltage_kV: float | None = None
    semiangle_mrad: float | None = None
    scan_sampling_A: float | tuple[float, float] | None = None
    source_path: str | None = None
    bf_center: tuple[float, float] | None = None
    bf_radius: float | None = None
    detected_bf_radius: float | None = None
    optuna_trials: list[dict] | None = None
    reused: bool = False
    saved_path: pathlib.Path | None = None
    metadata: dict[str, object] = field(default_factory=dict)

    @property
    def phase(self):
        """``angle(object_wave)``."""
        if _is_torch(self.object_wave):
            return self.object_wave.angle()
        return np.angle(self.object_wave)

    @property
    def amplitude(self):
        """``abs(object_wave)``."""
        if _is_torch(self.object_wave):
            return self.object_wave.abs()
        return np.abs(self.object_wave)


@runtime_checkable
class SSBProtocol(Protocol):
    """The backend contract of ``protocol.py:61-125``."""

    backend: Backend
    precision: SSBPrecision

    @property
    def scan_shape(self) -> tuple[int, int]: ...

    @property
    def detector_shape(self) -> tuple[int, int]: ...

    @property
    def num_bf(self) -> int: ...

    def fit(self, *, trials: int, refinement: str | None,
            search_ranges: dict[str, tuple[float, float] | float] | None,
            refine_lock: list[str] | None, seed: int, verbose: bool) -> SSBResult: ...

    def reconstruct_result(self, aberrations: dict[str, float], *,
                           compute_loss: bool = True) -> SSBResult: ...

    def cache_rotation(self, rotation_rad: float) -> None: ...

    def preview(self, aberrations: dict[str, float], *, compute_loss: bool,
                higher_order_magnitudes: np.ndarray | None,
                higher_order_angles: np.ndarray | None) -> tuple[np.ndarray, float | None]: ...

    def preview_context(self, num_bf: int) -> AbstractContextManager | None: ...

    def browser_state(self) -> SSBExportState: ...

    def export_brightfield(self, data, path_stem: str | pathlib.Path
                           ) -> tuple[pathlib.Path, float] | None: ...

    def close(self) -> None: ...


__all__ = ["SSBPrecision", "BrightfieldDisk", "SSBExportState", "SSBResult", "SSBProtocol"]
