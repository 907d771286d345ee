"""``HipSSBBackend``: the MI355X SSB session behind the reference's backend
contract.

Mirrors ``CudaSSBBackend`` (``cuda/backend.py:88-1420``): the same
constructor arguments, properties and methods, the same defaults, units
(aberrations in the reference's numeric convention, ``phi12`` in radians,
public rotation in degrees) and error types.  All computation runs in the
native library through ``HipSSBEngine``; this class only sequences calls,
keeps the session state and builds results.  There is no CPU fallback.
"""
from __future__ import annotations

import contextlib
import math
import pathlib
import time
from typing import Literal

import numpy as np

from ...types import BrightfieldDisk, SSBExportState, SSBPrecision, SSBResult
from . import optimizer as _opt
from .engine import HipSSBEngine, PlanSpec, SSBPlan, normalize_scan, scan_target


def _bf_subset_indices(full_num_bf: int, ratio: float) -> np.ndarray:
    """Uniform-stride BF subset for ``bf_subsample`` (``cuda/backend.py:21-43``):
    every round(1/ratio)-th pixel."""
    if ratio <= 0.0:
        raise ValueError(f"bf_subsample must be in (0, 1], got {ratio}")
    stride = max(1, int(round(1.0 / float(ratio))))
    return np.arange(0, full_num_bf, stride, dtype=np.int64)


class _PreviewSubset:
    """Prepared reduced-BF preview state; ``with`` routes previews to it."""

    def __init__(self, ssb: "HipSSBBackend", num_bf: int) -> None:
        full = ssb.num_bf
        count = max(1, min(int(num_bf), full))
        step = max(1, full // count)
        self._ssb = ssb
        self._engine = ssb._engine.subset(np.arange(0, full, step, dtype=np.int64)[:count])
        self._active = False

    @property
    def num_bf(self) -> int:
        return self._engine.num_bf

    def __enter__(self) -> "_PreviewSubset":
        if self._active or self._ssb._subset_engine is not None:
            raise RuntimeError("The prepared SSB BF subset is already active.")
        self._ssb._subset_engine = self._engine
        self._active = True
        return self

    def __exit__(self, *exc) -> None:
        self._ssb._subset_engine = None
        self._active = False


def _check_objective_mode(mode) -> None:
    """Objective modes of ``cuda/backend.py:1021-1026``.  The reference's
    "native" mode is a sparse-row objective with its own landscape
    (``cuda/kernels/fft256.py:343-350``); only the exact objective is built."""
    if mode not in {"native", "exact"}:
        raise ValueError(f"objective_mode must be 'native' or 'exact'; got {mode!r}.")
    if mode == "native":
        raise NotImplementedError("objective_mode='native' (the sparse-row objective) is not supported by the "
                                  "hip SSB backend; use objective_mode='exact'.")


def _wavelength_angstrom(energy_eV: float) -> float:
    from ....optics import electron_wavelength_angstrom
    return electron_wavelength_angstrom(energy_eV)


class HipSSBBackend:
    """SSB reconstruction on one MI355X (see module docstring)."""

    _DEFAULT_OPTIMIZE_RANGES = dict(_opt.DEFAULT_RANGES)
    backend = "hip"
    precision = SSBPrecision()

    def __init__(self, data, semiangle: float, scan_sampling: float | tuple[float, float],
                 det_sampling: float | tuple[float, float] | None = None, *, voltage_kV: float | None = None,
                 energy: float | None = None, scan_shape: tuple[int, int] | None = None,
                 bf_intensity_threshold: float = 0.0, bf_radius: float | None = None,
                 aberrations: dict[str, float] | None = None, rotation_angle_deg: float = 0.0,
                 detector_gain=None, gqk_storage: Literal["herm"] = "herm", device=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError("The hip SSB backend requires a ROCm GPU; no CPU fallback exists.")
        if voltage_kV is not None and energy is not None:
            raise ValueError("Specify voltage_kV or energy, not both.")
        if voltage_kV is not None:
            energy = voltage_kV * 1e3
        elif energy is None:
            raise ValueError("Specify voltage_kV (kV) or energy (eV).")
        if gqk_storage != "herm":
            raise ValueError(
                "gqk_storage='full' was removed from the SSB runtime path. Use the default exact "
                "Hermitian storage; low-level tests may construct canonical full-plane references directly.")
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        data = self._as_device_data(data, dev)
        if data.ndim == 3:
            if scan_shape is None:
                n = int(data.shape[0])
                side = int(n ** 0.5)
                if side * side != n:
                    raise ValueError(f"scan_shape is required: {n} frames is not a perfect square. "
                                     f"Pass scan_shape=(rows, cols).")
                scan_shape = (side, side)
            if scan_shape[0] * scan_shape[1] != data.shape[0]:
                raise ValueError("scan_shape does not match number of frames.")
            data = data.reshape(scan_shape[0], scan_shape[1], data.shape[1], data.shape[2])
        elif data.ndim != 4:
            raise ValueError("data must be 3D or 4D.")
        gain = None
        if detector_gain is not None:
            gain = torch.as_tensor(np.asarray(detector_gain.cpu() if isinstance(detector_gain, torch.Tensor)
                                              else detector_gain, dtype=np.float32)).to(dev)
            if tuple(gain.shape) != tuple(data.shape[2:]):
                raise ValueError("detector_gain shape must match detector shape; "
                                 f"got {tuple(gain.shape)}, expected {tuple(data.shape[2:])}.")
        self._source_shape = tuple(int(v) for v in data.shape)
        data = normalize_scan(data.contiguous())

        if isinstance(scan_sampling, (int, float)):
            scan_sampling = (float(scan_sampling), float(scan_sampling))
        wavelength = _wavelength_angstrom(energy)
        det_gpts = (int(data.shape[2]), int(data.shape[3]))
        scan_gpts = (int(data.shape[0]), int(data.shape[1]))
        if scan_gpts[0] not in (128, 256, 512, 1024):
            raise NotImplementedError(f"scan size {scan_gpts[0]} is not supported by the hip SSB backend; "
                                      "use 128, 256, 512 or 1024")
        self._detected_bf_radius: float | None = None
        if det_sampling is None:
            radius = self._detect_radius(data, dev, det_gpts, scan_gpts, wavelength, gain)
            self._detected_bf_radius = float(radius)
            det_sampling = (2 * semiangle) / radius
            det_sampling = (det_sampling, det_sampling)
        elif isinstance(det_sampling, (int, float)):
            det_sampling = (float(det_sampling), float(det_sampling))
        if aberrations is None:
            aberrations = {"C10": 0.0, "C12": 0.0, "phi12": 0.0}

        self.energy = energy
        self.voltage_kV = voltage_kV
        self.semiangle_mrad = semiangle
        self.semiangle_cutoff = semiangle
        self.scan_sampling = tuple(scan_sampling)
        self.angular_sampling = tuple(det_sampling)
        self.bf_intensity_threshold = float(bf_intensity_threshold)
        self.aberrations = {"C10": 0.0, "C12": 0.0, "phi12": 0.0, **aberrations}
        self._rotation_angle_rad = math.radians(rotation_angle_deg)
        self.gqk_storage = gqk_storage
        self.gpts = det_gpts
        self.wavelength = wavelength
        recip = (det_sampling[0] * 1e-3 / wavelength, det_sampling[1] * 1e-3 / wavelength)
        self.sampling = (1.0 / (recip[0] * det_gpts[0]), 1.0 / (recip[1] * det_gpts[1]))
        self._scan_shape = scan_gpts

        spec = PlanSpec(scan_gpts[0], det_gpts[0], det_gpts[1], det_gpts[0] * det_gpts[1], wavelength,
                        semiangle * 1e-3, det_sampling[0] * 1e-3, det_sampling[1] * 1e-3,
                        self.scan_sampling[0], self.scan_sampling[1])
        self._plan = SSBPlan(spec, device=dev)
        self._engine = HipSSBEngine(self._plan)
        self.detector_gain = gain
        if gain is not None:
            self._engine.set_detector_gain(gain)
        self._engine.mean_dp(data)
        bf = self._engine.select_bf(self.bf_intensity_threshold, bf_radius)
        self._engine.prepare_g(data)
        self._engine.cache_rotation(self._rotation_angle_rad)
        del data
        self.bf_inds_row = bf.rows
        self.bf_inds_col = bf.cols
        self.bf_center = bf.center
        self.dc_value = bf.dc
        self._bf_radius_arg = bf_radius

        self._best_loss = float("inf")
        self._elapsed_optimize = 0.0
        self._elapsed_grid = 0.0
        self._elapsed_refine = 0.0
        self._refine_method: str | None = None
        self._refine_nfev: int | None = None
        self._n_trials: int | None = None
        self._optuna_trials: list[dict] = []
        self._optimizer_objective_mode = "exact"
        self._subset_engine: HipSSBEngine | None = None
        self._bf_source_path: pathlib.Path | None = None
        self._bf_source_dtype: np.dtype | None = None
        self._bf_source_max_value: int | None = None

    # ------------------------------------------------------------- helpers
    @staticmethod
    def _as_device_data(data, dev):
        """Raw frames on the GPU in a sample type the native library reads:
        uint8/16/32, int8/16/32 and float32 as they are (the reference keeps
        the source dtype, ``cuda/backend.py:195-201``); bool as uint8;
        float16/float64 as float32 (the reference's float64 mean DP then sums
        the float32-rounded samples)."""
        import torch
        from .engine import dtype_code
        if not isinstance(data, torch.Tensor):
            arr = np.asarray(data)
            if arr.dtype == np.bool_:
                arr = arr.astype(np.uint8)
            elif arr.dtype in (np.float16, np.float64):
                arr = arr.astype(np.float32)
            data = torch.from_numpy(np.ascontiguousarray(arr))
        elif data.dtype == torch.bool:
            data = data.to(torch.uint8)
        elif data.dtype in (torch.float16, torch.bfloat16, torch.float64):
            data = data.to(dev).to(torch.float32)
        dtype_code(data.dtype)
        return data.to(dev)

    @staticmethod
    def _detect_radius(data, dev, det_gpts, scan_gpts, wavelength, gain=None) -> int:
        spec = PlanSpec(scan_gpts[0], det_gpts[0], det_gpts[1], 1, wavelength, 1e-3, 1e-3, 1e-3, 1.0, 1.0)
        eng = HipSSBEngine(SSBPlan(spec, device=dev))
        if gain is not None:
            eng.set_detector_gain(gain)
        eng.mean_dp(data)
        (_, _), radius = eng.detect_bf_radius()
        eng.plan.close()
        return radius

    def _resolve_coefs(self, C10, C12, phi12):
        if C10 is None:
            C10 = self.aberrations.get("C10", 0.0)
        if C12 is None:
            C12 = self.aberrations.get("C12", 0.0)
        if phi12 is None:
            phi12 = self.aberrations.get("phi12", 0.0)
        return C10, C12, phi12

    def _prepare(self, rotation_angle_rad=None) -> HipSSBEngine:
        if self._engine is None:
            raise ValueError("this SSB session is closed")
        if rotation_angle_rad is None:
            rotation_angle_rad = self._rotation_angle_rad
        eng = self._subset_engine if self._subset_engine is not None else self._engine
        eng.cache_rotation(float(rotation_angle_rad))
        return eng

    @contextlib.contextmanager
    def _bf_subset(self, indices):
        """Route every evaluation inside the block to a BF subset."""
        if self._subset_engine is not None:
            raise RuntimeError("A BF subset is already active.")
        self._subset_engine = self._engine.subset(indices)
        try:
            yield self._subset_engine
        finally:
            sub, self._subset_engine = self._subset_engine, None
            sub.plan.close()

    # ---------------------------------------------------------- properties
    @property
    def scan_shape(self) -> tuple[int, int]:
        return tuple(int(v) for v in self._scan_shape)

    @property
    def detector_shape(self) -> tuple[int, int]:
        return tuple(int(v) for v in self.gpts)

    @property
    def num_bf(self) -> int:
        return int(self.bf_inds_row.size)

    @property
    def G_qk(self):
        """Resident Hermitian Fourier stack ``[b, q_row, 0..N/2]`` (complex64, GPU)."""
        return self._engine.G_qk

    # ------------------------------------------------------ reconstruction
    def cache_rotation(self, rotation_rad: float) -> None:
        self._rotation_angle_rad = float(rotation_rad)
        self._prepare()

    def reconstruct(self, c10: float, c12: float, phi12: float):
        return self._prepare().reconstruct(c10, c12, phi12)

    def reconstruct_with_loss(self, c10: float, c12: float, phi12: float):
        return self._prepare().reconstruct_with_loss(c10, c12, phi12)

    def reconstruct_full(self, mags_m, angles_rad):
        return self._prepare().reconstruct_full(mags_m, angles_rad)

    def reconstruct_full_with_loss(self, mags_m, angles_rad):
        return self._prepare().reconstruct_full_with_loss(mags_m, angles_rad)

    @staticmethod
    def phase_to_numpy(phase) -> np.ndarray:
        import torch
        if isinstance(phase, torch.Tensor):
            return phase.detach().cpu().numpy().astype(np.float32, copy=False)
        return np.asarray(phase, dtype=np.float32)

    def reconstruct_phase(self, C10=None, C12=None, phi12=None, rotation_angle_rad=None):
        """Mean per-BF phase image (``cuda/backend.py:719-728``)."""
        C10, C12, phi12 = self._resolve_coefs(C10, C12, phi12)
        return self._prepare(rotation_angle_rad).reconstruct(C10, C12, phi12)

    def reconstruct_object(self, C10=None, C12=None, phi12=None, rotation_angle_rad=None):
        """Complex object of the Fourier-sum path (``cuda/backend.py:929-942``)."""
        C10, C12, phi12 = self._resolve_coefs(C10, C12, phi12)
        return self._prepare(rotation_angle_rad).reconstruct_object(C10, C12, phi12)

    def _ho_arrays_from_aberrations(self):
        """(mags[14], angles_rad[14], any_higher_order) of ``cuda/backend.py:730-769``."""
        a = self.aberrations
        mags = np.zeros(14, dtype=np.float64)
        angs = np.zeros(14, dtype=np.float64)
        mags[0] = float(a.get("C10", 0.0))
        mags[1] = float(a.get("C12", 0.0))
        angs[1] = float(a.get("phi12", 0.0))
        layout = [("C21", 2, True), ("C23", 3, True), ("C30", 4, False), ("C32", 5, True), ("C34", 6, True),
                  ("C41", 7, True), ("C43", 8, True), ("C45", 9, True), ("C50", 10, False),
                  ("C52", 11, True), ("C54", 12, True), ("C56", 13, True)]
        active = False
        for name, idx, has_angle in layout:
            if has_angle:
                mag = float(a.get(f"{name}_mag", 0.0))
                ang = float(a.get(f"{name}_angle", 0.0))
            else:
                mag, ang = float(a.get(name, 0.0)), 0.0
            if mag != 0.0:
                active = True
            mags[idx] = mag
            angs[idx] = math.radians(ang)
        return mags, angs, active

    def _compute_loss(self) -> float:
        return float(self._prepare().reconstruct_with_loss(
            self.aberrations["C10"], self.aberrations["C12"], self.aberrations["phi12"])[1])

    def result(self, *, compute_loss: bool = True) -> SSBResult:
        """Reconstruct with the current aberrations (``cuda/backend.py:771-851``)."""
        import torch
        mags, angs, ho_active = self._ho_arrays_from_aberrations()
        try:
            if ho_active:
                phase = self._prepare().reconstruct_full(mags, angs)
                obj = torch.polar(torch.ones_like(phase), phase).to(torch.complex64)
            else:
                obj = self.reconstruct_object()
        except torch.cuda.OutOfMemoryError:
            free = torch.cuda.mem_get_info(self._plan.device)[0] / 1e9
            raise MemoryError(
                f"Out of GPU VRAM during SSB reconstruction ({self.num_bf} BF pixels, {free:.1f} GB free).\n"
                "Try: SSB(..., bf_radius=<smaller>) to reduce BF pixel count, "
                "or restart the kernel to free stale GPU memory.") from None
        loss = None
        if compute_loss and not ho_active:
            try:
                loss = self._compute_loss()
            except (ValueError, ZeroDivisionError, FloatingPointError):
                loss = None
        elapsed = self._elapsed_optimize + self._elapsed_grid + self._elapsed_refine
        bf = self.browser_state().brightfield
        return SSBResult(
            object_wave=obj, backend="hip", aberrations=self.aberrations.copy(),
            rotation_angle_deg=math.degrees(self._rotation_angle_rad), loss=loss,
            elapsed=elapsed if elapsed > 0 else None, n_trials=self._n_trials, num_bf=self.num_bf,
            refine_method=self._refine_method, refine_nfev=self._refine_nfev,
            refine_elapsed=self._elapsed_refine if self._elapsed_refine > 0 else None,
            voltage_kV=self.voltage_kV, semiangle_mrad=self.semiangle_mrad,
            scan_sampling_A=self.scan_sampling[0], bf_center=bf.center_row_col, bf_radius=bf.radius_px,
            detected_bf_radius=bf.detected_radius_px, optuna_trials=self._optuna_trials)

    def reconstruct_result(self, aberrations: dict[str, float], *, compute_loss: bool = True) -> SSBResult:
        self.aberrations.update(aberrations)
        return self.result(compute_loss=compute_loss)

    def preview(self, aberrations: dict[str, float], *, compute_loss: bool,
                higher_order_magnitudes: np.ndarray | None, higher_order_angles: np.ndarray | None):
        """One transient float32 phase and optional exact loss (``cuda/backend.py:892-922``)."""
        if higher_order_magnitudes is not None:
            if compute_loss:
                phase, loss = self.reconstruct_full_with_loss(higher_order_magnitudes, higher_order_angles)
            else:
                phase, loss = self.reconstruct_full(higher_order_magnitudes, higher_order_angles), None
        elif compute_loss:
            phase, loss = self.reconstruct_with_loss(aberrations["C10"], aberrations["C12"],
                                                     aberrations["phi12"])
        else:
            phase, loss = self.reconstruct(aberrations["C10"], aberrations["C12"], aberrations["phi12"]), None
        return self.phase_to_numpy(phase), None if loss is None else float(loss)

    def preview_context(self, num_bf: int) -> "_PreviewSubset":
        """A reusable reduced-BF preview context (``cuda/backend.py:683-686``;
        ``_PreparedCudaBfSubset``, ``cuda/engine.py:306-380``): inside
        ``with``, previews use ``num_bf`` BF pixels taken at a uniform stride
        (every ``full // num_bf``-th, the first ``num_bf`` of them)."""
        return _PreviewSubset(self, num_bf)

    def browser_state(self) -> SSBExportState:
        """Compact host state (``cuda/engine.py:529-575``)."""
        eng = self._prepare()
        kx, ky, ak = (t.cpu().numpy() for t in eng.geometry())
        rows = np.asarray(self.bf_inds_row, dtype=np.int32)
        cols = np.asarray(self.bf_inds_col, dtype=np.int32)
        center = (float(self.bf_center[0]), float(self.bf_center[1]))
        dist = np.sqrt((rows - center[0]) ** 2 + (cols - center[1]) ** 2)
        radius = float(dist.max()) + 1e-3
        sampling = 0.5 * (float(self.angular_sampling[0]) + float(self.angular_sampling[1]))
        detected = 2.0 * float(self.semiangle_cutoff) / sampling
        disk = BrightfieldDisk(rows, cols, center, radius, detected, self.detector_shape)
        k = np.sqrt(kx.astype(np.float64) ** 2 + ky.astype(np.float64) ** 2)
        phi = np.arctan2(ky, kx)
        n = self._scan_shape[0]
        qx = np.fft.fftfreq(n, self.scan_sampling[0]).astype(np.float32)
        qy = np.fft.fftfreq(n, self.scan_sampling[1]).astype(np.float32)
        return SSBExportState(
            backend="hip", scan_shape=self.scan_shape, brightfield=disk, kx_bf=kx, ky_bf=ky, qx_1d=qx,
            qy_1d=qy, aperture_k=ak, alpha_k2=((k * self.wavelength) ** 2).astype(np.float32),
            cos2phi_k=np.cos(2 * phi).astype(np.float32), sin2phi_k=np.sin(2 * phi).astype(np.float32),
            wavelength_A=float(self.wavelength), semiangle_rad=float(self.semiangle_cutoff) * 1e-3,
            angular_sampling_rad=(float(self.angular_sampling[0]) * 1e-3, float(self.angular_sampling[1]) * 1e-3),
            sampling_A=(float(self.sampling[0]), float(self.sampling[1])), dc_value=complex(self.dc_value),
            bf_source_path=self._bf_source_path, bf_source_dtype=self._bf_source_dtype,
            bf_source_max_value=self._bf_source_max_value)

    def export_brightfield(self, data, path_stem):
        """Write exact detector counts as detector-major BF columns ``[bf, scan]``
        (``.u8`` when every count fits, else ``.u16``), as
        ``SSBEngine.write_exact_bf_source`` (``cuda/engine.py:577-701``);
        returns (path, seconds)."""
        import torch
        t0 = time.perf_counter()
        eng = self._prepare()
        if not isinstance(data, torch.Tensor):
            data = torch.from_numpy(np.ascontiguousarray(np.asarray(data)))
        if data.dtype not in (torch.uint8, torch.uint16):
            raise TypeError(f"Exact BF-source export supports uint8 or uint16 counts; got {data.dtype}.")
        if data.ndim != 4 or tuple(int(v) for v in data.shape[:2]) != self.scan_shape:
            raise ValueError("BF-source data must match the engine scan shape; "
                             f"got {tuple(data.shape)}, expected {self.scan_shape} + detector.")
        data = data.to(self._plan.device).contiguous()
        # uint8 counts always fit the narrow companion; their maximum is metadata only
        max_value = int(data.max().item()) if data.dtype == torch.uint8 else eng.max_count(data)
        narrow = max_value <= 255
        out_dtype = np.dtype(np.uint8 if narrow else np.uint16)
        path = pathlib.Path(path_stem).with_suffix(".u8" if narrow else ".u16").expanduser().resolve()
        path.parent.mkdir(parents=True, exist_ok=True)
        path.unlink(missing_ok=True)
        chunk = min(self.num_bf, 256)
        with path.open("wb") as fh:
            for b0 in range(0, self.num_bf, chunk):
                cnt = min(chunk, self.num_bf - b0)
                cols = eng.bf_columns(data, b0, cnt, narrow)
                fh.write(cols.cpu().numpy().astype(out_dtype, copy=False).tobytes())
        self._bf_source_path = path
        self._bf_source_dtype = out_dtype
        self._bf_source_max_value = max_value
        return path, time.perf_counter() - t0

    # -------------------------------------------------------- optimisation
    def _evaluate(self, points: np.ndarray) -> np.ndarray:
        """Exact loss of K (C10, C12, phi12) points at the session rotation and
        higher-order terms, one synchronisation."""
        pts = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        mags, angs, _ = self._ho_arrays_from_aberrations()
        m = np.repeat(mags[None, :], len(pts), axis=0)
        a = np.repeat(angs[None, :], len(pts), axis=0)
        # float32 parameters, as the reference evaluates its candidates
        m[:, 0] = pts[:, 0].astype(np.float32)
        m[:, 1] = pts[:, 1].astype(np.float32)
        a[:, 1] = pts[:, 2].astype(np.float32)
        losses = self._prepare().batch_phase_loss(m, a)
        return losses.cpu().numpy().astype(np.float64)

    def _evaluate_params(self, params: list[dict]) -> np.ndarray:
        """Exact loss of search candidates (``C10_nm``, ``C12_nm``, ``phi12_deg``,
        ``rotation_angle_deg``, ``C30``), one synchronisation."""
        return self._enqueue_params(params).cpu().numpy().astype(np.float64)

    def _enqueue_params(self, params: list[dict]):
        """The same losses as a device tensor, enqueued without synchronising."""
        mags, angs, _ = self._ho_arrays_from_aberrations()
        m = np.repeat(mags[None, :], len(params), axis=0)
        a = np.repeat(angs[None, :], len(params), axis=0)
        rot = np.empty(len(params), dtype=np.float64)
        for i, p in enumerate(params):
            m[i, 0] = np.float32(p["C10_nm"])
            m[i, 1] = np.float32(p["C12_nm"])
            a[i, 1] = np.float32(math.radians(p["phi12_deg"]))
            m[i, 4] = np.float32(p["C30"])
            rot[i] = math.radians(p["rotation_angle_deg"])
        return self._prepare().batch_phase_loss(m, a, rot)

    def optimize(self, aberrations=None, rotation_angle_deg=None, n_trials: int = 200, seed: int = 42,
                 verbose: bool = True, bf_subsample: float | None = None,
                 objective_mode: Literal["native", "exact"] = "exact"):
        """Global search (``cuda/backend.py:958-1110``): a seeded batched
        quasi-random design on the exact objective replaces Optuna TPE.

        Besides the reference's keys (``C10_nm``, ``C12_nm``, ``phi12_deg``
        and a ``rotation_angle_deg`` range), a ``C30`` range (Angstrom) may be
        searched; all candidates of a batch are evaluated with one launch
        sequence and one synchronisation.
        """
        t0 = time.perf_counter()
        _check_objective_mode(objective_mode)
        if bf_subsample is not None and bf_subsample < 1.0:
            with self._bf_subset(_bf_subset_indices(self.num_bf, bf_subsample)):
                return self.optimize(aberrations, rotation_angle_deg, n_trials, seed, verbose, None,
                                     objective_mode)
        if aberrations is None:
            aberrations = dict(self._DEFAULT_OPTIMIZE_RANGES)
        spec = dict(aberrations)
        if isinstance(rotation_angle_deg, tuple):
            spec["rotation_angle_deg"] = rotation_angle_deg
        elif rotation_angle_deg is not None:
            self._rotation_angle_rad = math.radians(float(rotation_angle_deg))
        self._optimizer_objective_mode = objective_mode
        defaults = {"C10_nm": float(self.aberrations.get("C10", 0.0)),
                    "C12_nm": float(self.aberrations.get("C12", 0.0)),
                    "phi12_deg": math.degrees(float(self.aberrations.get("phi12", 0.0))),
                    "C30": float(self.aberrations.get("C30", 0.0)),
                    "rotation_angle_deg": math.degrees(self._rotation_angle_rad)}
        keys = _opt.PARAM_KEYS + ("C30", "rotation_angle_deg")
        best_params, best_value, history = _opt.batched_search(self._evaluate_params, spec, defaults,
                                                               int(n_trials), seed=int(seed), keys=keys)
        self._optuna_trials = history
        self._best_loss = best_value
        for key, abr, conv in (("C10_nm", "C10", None), ("C12_nm", "C12", None),
                               ("phi12_deg", "phi12", math.radians), ("C30", "C30", None)):
            if key in best_params:
                v = best_params[key]
            elif key in aberrations and not isinstance(aberrations[key], tuple):
                v = aberrations[key]
            else:
                continue
            self.aberrations[abr] = conv(v) if conv else v
        if "rotation_angle_deg" in best_params:
            self._rotation_angle_rad = math.radians(best_params["rotation_angle_deg"])
        self._prepare()
        self._elapsed_optimize = time.perf_counter() - t0
        self._n_trials = int(n_trials)
        if verbose:
            self._print_summary("Optimize", self._elapsed_optimize)
        return self

    def refine(self, verbose: bool = True, xatol: float = 0.1, fatol: float = 1e-8, lock=None,
               bf_subsample: float | None = None, objective_mode=None):
        """Nelder-Mead on the exact objective (``cuda/backend.py:1112-1259``)."""
        t0 = time.perf_counter()
        if objective_mode is None:
            objective_mode = self._optimizer_objective_mode
        _check_objective_mode(objective_mode)
        if bf_subsample is not None and bf_subsample < 1.0:
            with self._bf_subset(_bf_subset_indices(self.num_bf, bf_subsample)):
                return self.refine(verbose, xatol, fatol, lock, None, objective_mode)
        self._optimizer_objective_mode = objective_mode
        lock = set(lock or [])
        keys = ["C10", "C12", "phi12"]
        free = [k for k in keys if k not in lock]
        if free != keys:
            raise ValueError(
                "GPU-only SSB Nelder-Mead does not support locked refinement yet "
                f"(locked: {', '.join(sorted(lock)) or 'unknown'}). Use no locked aberrations with "
                "refinement='nelder-mead', locked SSB reference mode, or add a GPU-batched locked refiner.")
        x0 = np.array([self.aberrations[k] for k in keys], dtype=np.float64)
        # the exact objective's tolerances of cuda/backend.py:1218-1230
        eff_xatol, eff_fatol, max_iter = xatol, fatol, 300
        if xatol == 0.1 and fatol == 1e-8:
            eff_xatol, eff_fatol, max_iter = 0.25, 2e-6, 160
        best_x, best_loss, n_evals = _opt.nelder_mead(self._evaluate, x0, eff_xatol, eff_fatol, max_iter)
        for i, k in enumerate(keys):
            self.aberrations[k] = float(best_x[i])
        self._best_loss = float(best_loss)
        self._elapsed_refine = time.perf_counter() - t0
        self._refine_method = "nelder-mead"
        self._refine_nfev = int(n_evals)
        if verbose:
            self._print_summary(f"Refine (nelder-mead, {n_evals} evals)", self._elapsed_refine)
        return self

    def _search_state(self) -> dict:
        """Current values in the search's parameter names."""
        return {"C10_nm": float(self.aberrations.get("C10", 0.0)), "C12_nm": float(self.aberrations.get("C12", 0.0)),
                "phi12_deg": math.degrees(float(self.aberrations.get("phi12", 0.0))),
                "C30": float(self.aberrations.get("C30", 0.0)),
                "rotation_angle_deg": math.degrees(self._rotation_angle_rad)}

    def _set_search_state(self, params: dict) -> None:
        for key, abr, conv in (("C10_nm", "C10", None), ("C12_nm", "C12", None),
                               ("phi12_deg", "phi12", math.radians), ("C30", "C30", None)):
            if key in params:
                self.aberrations[abr] = conv(params[key]) if conv else float(params[key])
        if "rotation_angle_deg" in params:
            self._rotation_angle_rad = math.radians(params["rotation_angle_deg"])
        self._prepare()

    def refine_keys(self, keys=("C10_nm", "C30", "rotation_angle_deg"), scales=(100.0, 1.0e5, 1.0),
                    xatol: float = 1e-3, fatol: float = 1e-9, max_iter: int = 200, verbose: bool = False):
        """Nelder-Mead (reference coefficients and tie rule) on the exact
        objective over any three search parameters, each divided by its scale;
        every candidate carries its own rotation and C30."""
        if len(keys) != 3 or len(scales) != 3:
            raise ValueError("refine_keys optimises exactly three parameters")
        t0 = time.perf_counter()
        base = self._search_state()

        def evaluate(points):
            params = []
            for x in np.asarray(points, dtype=np.float64).reshape(-1, 3):
                p = dict(base)
                for k, sc, v in zip(keys, scales, x):
                    p[k] = float(v) * sc
                params.append(p)
            return self._evaluate_params(params)

        x0 = np.array([base[k] / sc for k, sc in zip(keys, scales)])
        best_x, best_loss, n_evals = _opt.nelder_mead(evaluate, x0, xatol, fatol, max_iter)
        self._set_search_state({k: float(v) * sc for k, sc, v in zip(keys, scales, best_x)})
        self._best_loss = float(best_loss)
        self._refine_method = "nelder-mead"
        self._refine_nfev = int(n_evals)
        self._elapsed_refine = time.perf_counter() - t0
        if verbose:
            self._print_summary(f"Refine (nelder-mead, {n_evals} evals)", self._elapsed_refine)
        return self

    def refine_quadratic(self, keys=("C10_nm", "C30", "rotation_angle_deg"), radius=(75.0, 7.5e4, 2.5),
                         rounds: int = 4, batch: int = 16, shrink: float = 0.3, bounds=None, seed: int = 0,
                         verbose: bool = False):
        """Batched trust-region refinement on the exact objective: ``rounds``
        batches of ``batch`` candidates (one launch sequence each) plus one
        model-minimiser evaluation per round.  ``radius`` is the first box's
        half-width per key; ``bounds`` clips candidates."""
        t0 = time.perf_counter()
        base = self._search_state()
        keys = tuple(keys)
        bounds = bounds or {}
        lo = np.array([bounds.get(k, (-math.inf, math.inf))[0] for k in keys])
        hi = np.array([bounds.get(k, (-math.inf, math.inf))[1] for k in keys])

        def evaluate(points):
            params = []
            for x in np.asarray(points, dtype=np.float64).reshape(-1, len(keys)):
                p = dict(base)
                p.update({k: float(v) for k, v in zip(keys, x)})
                params.append(p)
            return self._evaluate_params(params)

        x0 = np.array([base[k] for k in keys])
        best_x, best_loss, n_evals = _opt.quadratic_refine(evaluate, x0, np.asarray(radius, dtype=np.float64),
                                                           rounds=rounds, batch=batch, shrink=shrink, seed=seed,
                                                           lo=lo, hi=hi)
        self._set_search_state({k: float(v) for k, v in zip(keys, best_x)})
        self._best_loss = float(best_loss)
        self._refine_method = "quadratic"
        self._refine_nfev = int(n_evals)
        self._elapsed_refine = time.perf_counter() - t0
        if verbose:
            self._print_summary(f"Refine (quadratic, {n_evals} evals)", self._elapsed_refine)
        return self

    def refine_newton(self, keys=("C10_nm", "C30", "rotation_angle_deg"), step=(10.0, 1.0e4, 0.25),
                      xtol=(0.05, 50.0, 2.0e-3), hmin=(1.0, 2500.0, 0.1), ftol: float = 1e-8, max_iter: int = 8,
                      polish: bool = True, bounds=None, verbose: bool = False):
        """Batched finite-difference Newton refinement on the exact objective
        (``optimizer.newton_refine``): one batch of 1 + 2d + d(d-1)/2
        candidates per iteration (10 for three keys), one synchronisation
        each.  ``step`` is the first stencil's half-width per key, ``hmin``
        its smallest (the exact loss moves in steps of ~3e-8 as BF pixels
        cross the aperture edge under rotation, so rotation differences
        below ~0.1 degree are noise), ``xtol`` the step size per key and
        ``ftol`` the predicted decrease below which it stops; ``polish`` adds
        one least-squares model minimiser at the end; ``bounds`` clips
        candidates."""
        t0 = time.perf_counter()
        base = self._search_state()
        keys = tuple(keys)
        bounds = bounds or {}
        lo = np.array([bounds.get(k, (-math.inf, math.inf))[0] for k in keys])
        hi = np.array([bounds.get(k, (-math.inf, math.inf))[1] for k in keys])

        def evaluate(points):
            params = []
            for x in np.asarray(points, dtype=np.float64).reshape(-1, len(keys)):
                p = dict(base)
                p.update({k: float(v) for k, v in zip(keys, x)})
                params.append(p)
            return self._evaluate_params(params)

        x0 = np.array([base[k] for k in keys])
        best_x, best_loss, n_evals = _opt.newton_refine(evaluate, x0, np.asarray(step, dtype=np.float64),
                                                        max_iter=max_iter, xtol=np.asarray(xtol, dtype=np.float64),
                                                        hmin=np.asarray(hmin, dtype=np.float64), lo=lo, hi=hi,
                                                        ftol=ftol, polish=polish)
        self._set_search_state({k: float(v) for k, v in zip(keys, best_x)})
        self._best_loss = float(best_loss)
        self._refine_method = "newton"
        self._refine_nfev = int(n_evals)
        self._elapsed_refine = time.perf_counter() - t0
        if verbose:
            self._print_summary(f"Refine (newton, {n_evals} evals)", self._elapsed_refine)
        return self

    _POLISH_BOX = {"C10_nm": 1.5, "C30": 6.0e3, "rotation_angle_deg": 0.1}

    def refine_polish(self, keys=("C10_nm", "C30", "rotation_angle_deg"), box=None, rounds: int = 2,
                      batch: int = 16, bounds=None, seed: int = 20, verbose: bool = False):
        """Least-squares quadratic polish on the exact objective
        (``optimizer.model_polish``): ``rounds`` batches of ``batch``
        candidates in a box of half-widths ``box`` (default ~3x the parameter
        change that moves the phase by 1e-4 rad RMS) around the current
        state, each followed by a move to the fitted model's minimiser; the
        state becomes the last minimiser (evaluated once more for
        ``_best_loss``).  Near the optimum the exact loss is rough at the
        ~3e-8 level, which makes its lowest sample a noisy estimate; the
        model's minimiser averages that roughness."""
        t0 = time.perf_counter()
        base = self._search_state()
        keys = tuple(keys)
        box = box or self._POLISH_BOX
        bounds = bounds or {}
        lo = np.array([bounds.get(k, (-math.inf, math.inf))[0] for k in keys])
        hi = np.array([bounds.get(k, (-math.inf, math.inf))[1] for k in keys])

        def evaluate(points):
            params = []
            for x in np.asarray(points, dtype=np.float64).reshape(-1, len(keys)):
                p = dict(base)
                p.update({k: float(v) for k, v in zip(keys, x)})
                params.append(p)
            return self._evaluate_params(params)

        x0 = np.array([base[k] for k in keys])
        x, n_evals = _opt.model_polish(evaluate, x0, np.array([box[k] for k in keys], dtype=np.float64),
                                       rounds=rounds, batch=batch, seed=seed, lo=lo, hi=hi)
        self._best_loss = float(np.asarray(evaluate(x[None, :])).reshape(-1)[0])
        self._set_search_state({k: float(v) for k, v in zip(keys, x)})
        # the polish completes a fit's refinement: its evaluations and time add to it
        self._refine_nfev = int(getattr(self, "_refine_nfev", 0) or 0) + int(n_evals) + 1
        self._refine_method = f"{self._refine_method}+polish" if self._refine_method else "polish"
        self._elapsed_refine = time.perf_counter() - t0 + float(getattr(self, "_elapsed_refine", 0.0) or 0.0)
        if verbose:
            self._print_summary(f"Polish (least-squares model, {n_evals + 1} evals)", time.perf_counter() - t0)
        return self

    def learned_start(self, net, feature_config) -> dict:
        """(C10, C30, rotation) predicted by the learned initializer from the
        resident Fourier stack, in the search's parameter names."""
        from ...learned import features_from_engine
        X = features_from_engine(self._prepare(), feature_config)
        c10, c30, rot = net.predict(X)[0]
        return {"C10_nm": float(c10), "C30": float(c30), "rotation_angle_deg": math.degrees(float(rot))}

    _LEARNED_HALF_WIDTHS = {"C10_nm": 75.0, "C30": 7.5e4, "rotation_angle_deg": 2.5}
    # first Newton stencils and stopping decreases: after the global search,
    # and after the learned start's local search (a box of the half-widths
    # above, 16 trials).  Measured over 13 held-out scan-256 datasets: the
    # learned start stops at a predicted decrease of 6e-8 (2x the loss's
    # noise) with no loss of phase error, the search's farther start needs
    # 1e-8 to keep it.
    _NEWTON_STEP = (10.0, 1.0e4, 0.25)
    _LEARNED_NEWTON_STEP = (5.0, 5.0e3, 0.15)
    _NEWTON_FTOL = 1e-8
    _LEARNED_NEWTON_FTOL = 6e-8
    _FIT_BOUNDS = {"C10_nm": (-500.0, 500.0), "C30": (-2.0e5, 2.0e5), "rotation_angle_deg": (-5.0, 5.0)}

    def _local_refine(self, refine: str, keys, radius, rounds, batch: int, bounds, seed: int,
                      xatol: float, verbose: bool, newton_step=None, newton_ftol: float = 1e-8) -> None:
        if refine == "nelder-mead":
            self.refine_keys(keys, xatol=xatol, fatol=1e-10, max_iter=400, verbose=verbose)
        elif refine == "quadratic":
            self.refine_quadratic(keys, radius=radius, rounds=rounds or 4, batch=batch, bounds=bounds, seed=seed,
                                  verbose=verbose)
        elif refine == "newton":
            self.refine_newton(keys, step=newton_step or self._NEWTON_STEP, ftol=newton_ftol, max_iter=rounds or 8,
                               bounds=bounds, verbose=verbose)
        elif refine is not None:
            raise ValueError(f"refine must be 'nelder-mead', 'quadratic', 'newton' or None; got {refine!r}")

    def fit_learned(self, net, feature_config, *, trials: int = 16, refine: str | None = "newton",
                    xatol: float = 1e-4, rounds: int | None = None, half_widths=None, bounds=None, seed: int = 42,
                    newton_step=None, newton_ftol: float | None = None, polish: bool = True,
                    verbose: bool = False):
        """Learned start, a ``trials``-candidate batched Sobol search in a box
        around it (clipped to ``bounds``, the global search box), then a local
        refinement of (C10, C30, rotation) on the exact objective:
        ``"newton"`` (default: batched finite-difference Newton,
        ``refine_newton``, at most ``rounds`` batches, default 8),
        ``"nelder-mead"`` (the reference's coefficients, ``refine_keys``) or
        ``"quadratic"`` (``refine_quadratic``, ``rounds`` default 4); after
        the Newton refine, with ``polish``, the least-squares model polish
        (``refine_polish``), whose minimiser is the result.
        The global batched search (``fit_search``) is the fallback when no
        trained network is available."""
        t0 = time.perf_counter()
        start = self.learned_start(net, feature_config)
        half = dict(self._LEARNED_HALF_WIDTHS)
        half.update(half_widths or {})
        bounds = bounds or dict(self._FIT_BOUNDS)
        spec = {k: v for k, v in self._search_state().items() if k not in start}
        for k, v in start.items():
            lo, hi = bounds.get(k, (-math.inf, math.inf))
            c = min(max(v, lo), hi)
            spec[k] = (max(lo, c - half[k]), min(hi, c + half[k]))
        if trials > 0:
            best, best_value, history = _opt.batched_search(self._evaluate_params, spec, self._search_state(),
                                                            trials, seed=seed, refine_fraction=0.0,
                                                            keys=_opt.PARAM_KEYS + ("C30", "rotation_angle_deg"))
            self._set_search_state(best)
            self._best_loss = best_value
            self._optuna_trials = history
        else:
            self._set_search_state({k: min(max(v, bounds.get(k, (-math.inf, math.inf))[0]),
                                           bounds.get(k, (-math.inf, math.inf))[1]) for k, v in start.items()})
        keys = tuple(start)
        self._local_refine(refine, keys, [half[k] / 2 for k in keys], rounds, 16, bounds, seed, xatol, verbose,
                           newton_step or self._LEARNED_NEWTON_STEP,
                           self._LEARNED_NEWTON_FTOL if newton_ftol is None else newton_ftol)
        if polish and refine == "newton":
            self.refine_polish(keys, bounds=bounds, verbose=verbose)
        self._learned_start = start
        self._n_trials = trials
        self._elapsed_optimize = time.perf_counter() - t0
        return self

    def fit_search(self, *, trials: int = 512, refine: str | None = "newton", xatol: float = 1e-4,
                   rounds: int | None = None, bounds=None, seed: int = 42, newton_step=None,
                   newton_ftol: float | None = None, polish: bool = True, verbose: bool = False):
        """Plain batched fit without the network: a ``trials``-candidate global
        Sobol search over ``bounds`` (C10, C30, rotation), then the same local
        refinement and polish as ``fit_learned`` from its best point (a wider
        first Newton stencil: the global search's best point is farther
        out)."""
        t0 = time.perf_counter()
        bounds = bounds or dict(self._FIT_BOUNDS)
        self.optimize(aberrations=dict(bounds, C12_nm=0.0, phi12_deg=0.0), n_trials=trials, seed=seed,
                      verbose=False)
        keys = tuple(bounds)
        self._local_refine(refine, keys, [0.05 * (hi - lo) for lo, hi in bounds.values()], rounds, 16, bounds,
                           seed, xatol, verbose, newton_step, self._NEWTON_FTOL if newton_ftol is None else newton_ftol)
        if polish and refine == "newton":
            self.refine_polish(keys, bounds=bounds, verbose=verbose)
        self._elapsed_optimize = time.perf_counter() - t0
        return self

    _DEFAULT_GRID_HALF_WIDTHS = {"C10_nm": 50, "C12_nm": 20, "phi12_deg": 30}
    _DEFAULT_GRID_POINTS = {"C10_nm": 21, "C12_nm": 11, "phi12_deg": 13}

    def grid_search(self, window: dict | None = None, n_points: dict | None = None, verbose: bool = True):
        """Exhaustive (C10, C12, phi12) grid around the current aberrations
        (``cuda/backend.py:1261-1374``): same default half-widths and point
        counts, C12 clipped at 0, first minimum (``argmin``) wins.  All points
        are evaluated in batches of exact losses, one synchronisation each."""
        from itertools import product
        t0 = time.perf_counter()
        window = window or {}
        n_points = n_points or {}
        hw = {k: window.get(k, v) for k, v in self._DEFAULT_GRID_HALF_WIDTHS.items()}
        npt = {k: n_points.get(k, v) for k, v in self._DEFAULT_GRID_POINTS.items()}
        c10c, c12c = self.aberrations["C10"], self.aberrations["C12"]
        phic = math.degrees(self.aberrations["phi12"])
        c10v = np.linspace(c10c - hw["C10_nm"], c10c + hw["C10_nm"], npt["C10_nm"])
        c12v = np.linspace(max(0, c12c - hw["C12_nm"]), c12c + hw["C12_nm"], npt["C12_nm"])
        phiv = [math.radians(v) for v in np.linspace(phic - hw["phi12_deg"], phic + hw["phi12_deg"],
                                                     npt["phi12_deg"])]
        params = np.array(list(product(c10v, c12v, phiv)), dtype=np.float64)
        if verbose:
            print(f"Grid search: {len(params)} combinations ({len(c10v)}x{len(c12v)}x{len(phiv)})")
        losses = np.concatenate([self._evaluate(params[s:s + 256]) for s in range(0, len(params), 256)])
        best = int(np.argmin(losses))
        self._best_loss = float(losses[best])
        self.aberrations["C10"], self.aberrations["C12"], self.aberrations["phi12"] = (float(v) for v in params[best])
        self._elapsed_grid = time.perf_counter() - t0
        if verbose:
            self._print_summary(f"Grid ({len(params)} evals)", self._elapsed_grid)
        return self

    def fit(self, *, trials: int, refinement: str | None, search_ranges, refine_lock, seed: int,
            verbose: bool, save_to=None, force: bool = False, source_path=None) -> SSBResult:
        """The shared fit contract (``cuda/backend.py:853-879``).

        ``save_to`` (with ``source_path``, the detector file the data came
        from) reuses an exact prior result when nothing that determines it
        changed, as ``SSB.fit`` does (``qg/ssb/workflow.py:1043-1103``);
        ``force`` recomputes."""
        def run():
            if trials:
                self.optimize(aberrations=search_ranges, n_trials=int(trials), seed=int(seed), verbose=verbose,
                              objective_mode="exact")
            if refinement == "nelder-mead":
                self.refine(verbose=verbose, lock=refine_lock, objective_mode="exact")
            return self.result()

        if save_to is None:
            return run()
        settings = {"trials": int(trials), "refinement": refinement, "search_ranges": search_ranges,
                    "refine_lock": refine_lock, "seed": int(seed),
                    "starting_aberrations": dict(self.aberrations),
                    "starting_rotation_deg": math.degrees(self._rotation_angle_rad)}
        return self._saved("fit", settings, run, save_to, force, source_path, verbose)

    def reconstruct_saved(self, aberrations: dict[str, float], *, save_to, source_path, compute_loss: bool = True,
                          force: bool = False, verbose: bool = False) -> SSBResult:
        """``reconstruct_result`` with the saved-result reuse of ``SSB.reconstruct``
        (``qg/ssb/workflow.py:1113-1170``)."""
        coefs = {**self.aberrations, **aberrations}
        settings = {"aberrations": coefs, "compute_loss": bool(compute_loss)}
        return self._saved("reconstruct", settings,
                           lambda: self.reconstruct_result(aberrations, compute_loss=compute_loss),
                           save_to, force, source_path, verbose)

    def _saved(self, operation, settings, run, save_to, force, source_path, verbose) -> SSBResult:
        from ...persistence import SCHEMA, json_value, path_signature, reuse_or_run, software_signature
        if source_path is None:
            raise ValueError("save_to requires source_path when SSB.from_array() is used. Pass the detector "
                             "source path so saved results cannot be reused for a different array accidentally.")
        if not pathlib.Path(source_path).expanduser().exists():
            raise FileNotFoundError(f"Cannot save an SSB result because source_path does not exist: "
                                    f"{pathlib.Path(source_path).expanduser()}")
        signature = json_value({
            "schema": SCHEMA, "software": software_signature(), "operation": operation,
            "source": {"detector": path_signature(source_path)},
            "data": {"dtype": "uint16", "shape": self._source_shape, "scan_shape": self.scan_shape},
            "instrument": {"energy_eV": self.energy, "semiangle_mrad": self.semiangle_mrad,
                           "scan_sampling_A": self.scan_sampling, "det_sampling": self.angular_sampling},
            "ssb": {"backend": self.backend, "rotation_angle_deg": math.degrees(self._rotation_angle_rad),
                    "bf_intensity_threshold": self.bf_intensity_threshold, "bf_radius": self._bf_radius_arg},
            "settings": settings,
        })
        result = reuse_or_run(operation, save_to, signature, run, force=force, device=self._plan.device,
                              verbose=verbose, input_metadata={"source_path": str(source_path)})
        if result.reused:
            self.aberrations.update({k: float(v) for k, v in result.aberrations.items()})
            self._rotation_angle_rad = math.radians(result.rotation_angle_deg)
            self._prepare()
        return result

    def _print_summary(self, stage: str, elapsed: float) -> None:
        a = self.aberrations
        print(f"  {stage}: loss={self._best_loss:.6f}  C10={a['C10']:.1f} nm  C12={a['C12']:.1f} nm  "
              f"phi12={math.degrees(a['phi12']):.1f}°  {elapsed:.1f}s")

    # ------------------------------------------------------------ lifetime
    def free(self) -> None:
        if getattr(self, "_plan", None) is not None:
            self._plan.close()
        self._plan = None
        self._engine = None

    def close(self) -> None:
        self.free()


__all__ = ["HipSSBBackend", "scan_target"]
