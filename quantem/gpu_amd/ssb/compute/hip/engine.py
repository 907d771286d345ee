"""HIP SSB engine: a plan on one MI355X plus the operations of the reference engine.

Mirrors the seam the reference's tests drive, ``SSBEngine``
(``cuda/engine.py:402-2295``): construction from a Fourier stack and BF
geometry, ``cache_rotation``, ``reconstruct_object``, ``reconstruct``,
``reconstruct_with_loss``, ``variance_loss(_batch)``, ``export_state``.
Device memory and streams come from PyTorch; every computation runs in the
native library through ``_abi``.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _abi


def _host(x):
    """A host copy of a small array given as NumPy, a list or a torch tensor."""
    try:
        import torch
        if isinstance(x, torch.Tensor):
            return x.detach().cpu().numpy()
    except ModuleNotFoundError:
        pass
    return x


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("The hip SSB backend requires a ROCm GPU; no CPU fallback exists.")
    return torch


def check_out(tensor, name: str, numel: int, dtype, device) -> None:
    """Reject a caller-supplied output the native call would write out of
    bounds or through a host address: it must be a contiguous ``dtype``
    tensor on ``device`` holding at least ``numel`` elements.  ``None``
    (no output wanted) passes."""
    if tensor is None:
        return
    import torch
    if not isinstance(tensor, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if tensor.dtype != dtype:
        raise ValueError(f"{name} must have dtype {dtype}; got {tensor.dtype}")
    if not tensor.is_cuda or tensor.device != torch.device(device):
        raise ValueError(f"{name} must live on {device}; got {tensor.device}")
    if not tensor.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if tensor.numel() < numel:
        raise ValueError(f"{name} holds {tensor.numel()} elements; {numel} are written")


# QSSB_DTYPE_* codes of the sample types the native library reads
_DTYPE_CODES = {"uint8": 1, "uint16": 2, "uint32": 3, "int8": 4, "int16": 5, "int32": 6, "float32": 7}


def dtype_code(dtype) -> int:
    """QSSB_DTYPE_* code of a torch sample dtype; NotImplementedError for others."""
    name = str(dtype).replace("torch.", "")
    if name not in _DTYPE_CODES:
        raise NotImplementedError(f"the hip SSB backend reads uint8, uint16, uint32, int8, int16, int32 or float32 "
                                  f"detector frames; got {dtype}")
    return _DTYPE_CODES[name]


def scan_target(rows: int, cols: int) -> int:
    """Supported square scan size for a rows x cols scan (``cuda/backend.py:238-246``)."""
    return int(_abi.lib().ssb_scan_target(int(rows), int(cols)))


def normalize_scan(data):
    """Crop/pad a resident ``[H, W, k_r, k_c]`` scan (any supported sample
    type) to the supported square size with the mean-frame padding of
    ``cuda/backend.py:229-269`` (returns ``data`` itself when it is already
    supported)."""
    torch = _torch()
    if not isinstance(data, torch.Tensor) or not data.is_cuda:
        raise ValueError("data must be a torch tensor on the GPU")
    code = dtype_code(data.dtype)
    if data.ndim != 4:
        raise ValueError("data must be 4D (scan_row, scan_col, k_row, k_col)")
    H, W, kr, kc = (int(v) for v in data.shape)
    T = scan_target(H, W)
    if (H, W) == (T, T):
        return data
    src = data.contiguous()
    out = torch.empty((T, T, kr, kc), dtype=data.dtype, device=data.device)
    with torch.cuda.device(data.device):
        stream = _abi.stream_handle(torch.cuda.current_stream(data.device))
        _abi.call("ssb_normalize_scan_typed", _abi.ptr(src), code, H, W, kr, kc, _abi.ptr(out), T, stream)
    return out


@dataclass(frozen=True)
class PlanSpec:
    scan_n: int
    det_rows: int
    det_cols: int
    max_bf: int
    wavelength_A: float
    semiangle_rad: float
    ang_row_rad: float
    ang_col_rad: float
    scan_sampling_row_A: float
    scan_sampling_col_A: float


class SSBPlan:
    """Owns one native plan (device buffers for one scan/detector geometry)."""

    def __init__(self, spec: PlanSpec, device=None):
        torch = _torch()
        self.spec = spec
        dev = torch.device(device if device is not None else "cuda")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _abi.call("ssb_plan_create", spec.scan_n, spec.det_rows, spec.det_cols, spec.max_bf,
                      spec.wavelength_A, spec.semiangle_rad, spec.ang_row_rad, spec.ang_col_rad,
                      spec.scan_sampling_row_A, spec.scan_sampling_col_A, ctypes.byref(handle))
        self._h = handle

    @property
    def handle(self) -> ctypes.c_void_p:
        if self._h is None:
            raise ValueError("SSB plan is closed")
        return self._h

    def close(self) -> None:
        if getattr(self, "_h", None) is not None:
            _abi.lib().ssb_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class BrightFieldState:
    rows: np.ndarray
    cols: np.ndarray
    center: tuple[float, float]
    radius2: float
    dc: complex


class HipSSBEngine:
    """SSB computations on one plan; arrays are torch tensors on the plan's GPU."""

    backend = "hip"

    def __init__(self, plan: SSBPlan):
        self.plan = plan
        self.torch = _torch()
        self._bf: BrightFieldState | None = None
        self._rotation: float | None = None
        self._has_g = False
        self._device_bf = False

    # ------------------------------------------------------------------ utils
    @property
    def n(self) -> int:
        return self.plan.spec.scan_n

    def _stream(self):
        return _abi.stream_handle(self.torch.cuda.current_stream(self.plan.device))

    def _check_data(self, data) -> int:
        """Validate resident raw frames; returns their QSSB_DTYPE_* code."""
        torch = self.torch
        if not isinstance(data, torch.Tensor):
            raise ValueError("data must be a torch tensor on the GPU")
        code = dtype_code(data.dtype)
        if not data.is_cuda or data.device != self.plan.device:
            raise ValueError(f"data must live on {self.plan.device}")
        if not data.is_contiguous():
            raise ValueError("data must be contiguous")
        s = self.plan.spec
        if data.ndim != 4 or tuple(data.shape) != (s.scan_n, s.scan_n, s.det_rows, s.det_cols):
            raise ValueError(
                f"data must have shape ({s.scan_n}, {s.scan_n}, {s.det_rows}, {s.det_cols}); "
                f"got {tuple(data.shape)}"
            )
        return code

    def _check_u16(self, data):
        if self._check_data(data) != _DTYPE_CODES["uint16"]:
            raise ValueError("data must be a torch.uint16 tensor on the GPU")

    # ------------------------------------------------------------ preparation
    def mean_dp(self, data):
        """(uint64 sums, float32 mean DP) of resident raw frames; the sums are
        None for float32 frames (summed in double-float).  The mean carries
        the detector gain when one is set."""
        torch = self.torch
        code = self._check_data(data)
        s = self.plan.spec
        sums = None
        if code != _DTYPE_CODES["float32"]:
            sums = torch.empty((s.det_rows, s.det_cols), dtype=torch.uint64, device=self.plan.device)
        mean = torch.empty((s.det_rows, s.det_cols), dtype=torch.float32, device=self.plan.device)
        _abi.call("ssb_mean_dp_typed", self.plan.handle, _abi.ptr(data), code, s.scan_n * s.scan_n,
                  _abi.ptr(sums), _abi.ptr(mean), self._stream())
        return sums, mean

    def set_detector_gain(self, gain) -> None:
        """Per-pixel detector gain ``[det_rows, det_cols]`` (float32; None
        clears it), applied as ``cuda/backend.py:220-226``, ``442-443``,
        ``574-576`` apply ``detector_gain``."""
        torch = self.torch
        s = self.plan.spec
        if gain is None:
            _abi.call("ssb_set_detector_gain", self.plan.handle, None, self._stream())
            return
        g = gain if isinstance(gain, torch.Tensor) else torch.as_tensor(np.asarray(gain, dtype=np.float32))
        g = g.to(device=self.plan.device, dtype=torch.float32).contiguous()
        if tuple(g.shape) != (s.det_rows, s.det_cols):
            raise ValueError("detector_gain shape must match detector shape; "
                             f"got {tuple(g.shape)}, expected {(s.det_rows, s.det_cols)}.")
        _abi.call("ssb_set_detector_gain", self.plan.handle, _abi.ptr(g), self._stream())
        self._gain = g             # keep the source alive until the copy has run

    def select_bf(self, threshold: float = 0.0, bf_radius: float | None = None) -> BrightFieldState:
        """Device BF selection from the last mean DP; reads the result back."""
        _abi.call("ssb_select_bf", self.plan.handle, float(threshold),
                  float(bf_radius) if bf_radius is not None else 0.0, self._stream())
        self._has_g = False
        self._rotation = None
        return self.bf_state()

    def bf_state(self) -> BrightFieldState:
        cap = self.plan.spec.max_bf
        rows = np.zeros(cap, dtype=np.int32)
        cols = np.zeros(cap, dtype=np.int32)
        nbf = ctypes.c_int(0)
        info = np.zeros(6, dtype=np.float32)
        _abi.call("ssb_get_bf", self.plan.handle,
                  rows.ctypes.data_as(_abi._ip), cols.ctypes.data_as(_abi._ip), cap,
                  ctypes.byref(nbf), info.ctypes.data_as(_abi._fp), self._stream())
        status = int(info[5])
        if status == 1:
            raise ValueError("No bright-field pixels found with the given threshold. Check that the "
                             "data contains a visible BF disk, or lower the threshold.")
        if status == 2:
            raise ValueError("No bright-field pixels within bf_radius. Increase bf_radius or check "
                             "detector geometry.")
        if status == 3:
            raise ValueError(f"More bright-field pixels than the plan capacity ({cap}).")
        k = int(nbf.value)
        self._bf = BrightFieldState(rows[:k].copy(), cols[:k].copy(),
                                    (float(info[0]), float(info[1])), float(info[2]),
                                    complex(float(info[3]), float(info[4])))
        return self._bf

    def detect_bf_radius(self, threshold_ratio: float = 0.1):
        """((centre row, centre col), radius) of the probe disk in the last
        mean DP (``qg/detector/compute/cuda/probe.py:12-123``)."""
        r, c, rad = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
        _abi.call("ssb_detect_bf_radius", self.plan.handle, float(threshold_ratio), ctypes.byref(r),
                  ctypes.byref(c), ctypes.byref(rad), self._stream())
        return (int(r.value), int(c.value)), int(rad.value)

    def max_count(self, data) -> int:
        """Largest count in resident uint16 data (synchronises)."""
        self._check_u16(data)
        out = ctypes.c_uint(0)
        _abi.call("ssb_max_u16", _abi.ptr(data), int(data.numel()), ctypes.byref(out), self._stream())
        return int(out.value)

    def bf_columns(self, data, bf_offset: int, count: int, narrow_u8: bool = False, out=None):
        """Exact detector-major BF columns ``[count, N*N]`` of uint16 frames
        (uint16, or uint8 when ``narrow_u8``) or uint8 frames (uint8)."""
        torch = self.torch
        code = self._check_data(data)
        if code not in (_DTYPE_CODES["uint8"], _DTYPE_CODES["uint16"]):
            raise TypeError(f"Exact BF-source export supports uint8 or uint16 counts; got {data.dtype}.")
        u8_in = code == _DTYPE_CODES["uint8"]
        dt = torch.uint8 if (narrow_u8 or u8_in) else torch.uint16
        if out is None:
            out = torch.empty((count, self.n * self.n), dtype=dt, device=self.plan.device)
        check_out(out, "out", int(count) * self.n * self.n, dt, self.plan.device)
        if u8_in:
            _abi.call("ssb_bf_columns_u8", self.plan.handle, _abi.ptr(data), int(bf_offset), int(count),
                      _abi.ptr(out), self._stream())
        else:
            _abi.call("ssb_bf_columns_u16", self.plan.handle, _abi.ptr(data), int(bf_offset), int(count),
                      1 if narrow_u8 else 0, _abi.ptr(out), self._stream())
        return out

    def set_bf(self, rows, cols, center, dc) -> None:
        rows = np.ascontiguousarray(np.asarray(rows, dtype=np.int32).reshape(-1))
        cols = np.ascontiguousarray(np.asarray(cols, dtype=np.int32).reshape(-1))
        if rows.shape != cols.shape or rows.size == 0:
            raise ValueError("bf rows and cols must be non-empty matching vectors")
        dc = complex(dc)
        _abi.call("ssb_set_bf", self.plan.handle, rows.ctypes.data_as(_abi._ip),
                  cols.ctypes.data_as(_abi._ip), int(rows.size), float(center[0]), float(center[1]),
                  float(dc.real), float(dc.imag), self._stream())
        self._bf = BrightFieldState(rows.copy(), cols.copy(), (float(center[0]), float(center[1])),
                                    0.0, dc)
        self._has_g = False
        self._rotation = None

    @property
    def num_bf(self) -> int:
        if self._bf is None:
            if not self._device_bf:
                raise ValueError("no bright-field selection")
            self.bf_state()
        return int(self._bf.rows.size)

    def subset(self, indices) -> "HipSSBEngine":
        """A new engine on BF pixels ``indices`` of this one's resident stack
        (same geometry and dc; its own plan), for reduced-BF previews and
        searches (``_PreparedCudaBfSubset``, ``use_bf_subset``:
        ``cuda/engine.py:306-380``, ``720-800``)."""
        from dataclasses import replace
        idx = np.asarray(indices, dtype=np.int64).reshape(-1)
        bf = self.bf_state() if self._bf is None else self._bf
        if idx.size == 0 or idx.min() < 0 or idx.max() >= bf.rows.size:
            raise ValueError("BF subset indices must be non-empty and inside the BF list")
        if not self._has_g:
            raise ValueError("no resident Fourier stack")
        g = self.G_qk[self.torch.as_tensor(idx, device=self.plan.device)]
        sub = HipSSBEngine(SSBPlan(replace(self.plan.spec, max_bf=int(idx.size)), device=self.plan.device))
        sub.set_bf(bf.rows[idx], bf.cols[idx], bf.center, bf.dc)
        sub.import_g(g)
        return sub

    def set_detector_sampling(self, recip_row: float, recip_col: float) -> None:
        _abi.call("ssb_set_detector_sampling", self.plan.handle, float(recip_row), float(recip_col))
        self._rotation = None

    def set_scan_frequencies(self, q_row, q_col) -> None:
        qr = np.ascontiguousarray(np.asarray(q_row, dtype=np.float32).reshape(-1))
        qc = np.ascontiguousarray(np.asarray(q_col, dtype=np.float32).reshape(-1))
        if qr.size != self.n or qc.size != self.n:
            raise ValueError(f"scan frequencies must have {self.n} entries")
        _abi.call("ssb_set_scan_frequencies", self.plan.handle, qr.ctypes.data_as(_abi._fp),
                  qc.ctypes.data_as(_abi._fp))

    def cache_rotation(self, rotation_rad: float, force: bool = False) -> None:
        if self._rotation == rotation_rad and not force:
            return
        _abi.call("ssb_cache_rotation", self.plan.handle, float(rotation_rad), self._stream())
        self._rotation = float(rotation_rad)

    def geometry(self):
        """(kx, ky, aperture) float32 tensors of the cached rotation."""
        torch = self.torch
        nb = self.num_bf
        out = [torch.empty(nb, dtype=torch.float32, device=self.plan.device) for _ in range(3)]
        _abi.call("ssb_get_geometry", self.plan.handle, *(_abi.ptr(t) for t in out), nb, self._stream())
        return tuple(out)

    def prepare_g(self, data) -> None:
        code = self._check_data(data)
        _abi.call("ssb_prepare_g_typed", self.plan.handle, _abi.ptr(data), code, self._stream())
        self._has_g = True

    def import_g(self, g_qk) -> None:
        torch = self.torch
        g = g_qk
        if not isinstance(g, torch.Tensor):
            g = torch.as_tensor(np.asarray(g, dtype=np.complex64))
        g = g.to(device=self.plan.device, dtype=torch.complex64).contiguous()
        if g.ndim != 3 or g.shape[1] != self.n or g.shape[2] not in (self.n, self.n // 2 + 1):
            raise ValueError(
                f"G_qk must have shape (num_bf, {self.n}, {self.n}) or Hermitian "
                f"(num_bf, {self.n}, {self.n // 2 + 1})"
            )
        if g.shape[0] != self.num_bf:
            raise ValueError(
                "G_qk first dimension must match bf_inds_row and bf_inds_col lengths; "
                f"got G_qk.shape[0]={g.shape[0]}, num_bf={self.num_bf}."
            )
        _abi.call("ssb_import_g", self.plan.handle, _abi.ptr(g), int(g.shape[0]), int(g.shape[2]),
                  self._stream())
        self._g_keepalive = g
        self.torch.cuda.current_stream(self.plan.device).synchronize()
        self._g_keepalive = None
        self._has_g = True

    @property
    def G_qk(self):
        """Resident Fourier stack in the reference layout ``[b, q_row, q_col<=N/2]``."""
        torch = self.torch
        out = torch.empty((self.num_bf, self.n, self.n // 2 + 1), dtype=torch.complex64,
                          device=self.plan.device)
        _abi.call("ssb_export_g", self.plan.handle, _abi.ptr(out), self.num_bf, self._stream())
        return out

    # -------------------------------------------------------- reconstruction
    def reconstruct_object(self, c10: float, c12: float, phi12: float, phase_out=None):
        """Complex object (complex64 [N, N]) from the resident stack."""
        torch = self.torch
        if not self._has_g:
            raise ValueError("no resident Fourier stack")
        check_out(phase_out, "phase_out", self.n * self.n, torch.float32, self.plan.device)
        if self._rotation is None:
            self.cache_rotation(0.0)
        obj = torch.empty((self.n, self.n), dtype=torch.complex64, device=self.plan.device)
        _abi.call("ssb_object", self.plan.handle, float(c10), float(c12), float(phi12),
                  _abi.ptr(obj), _abi.ptr(phase_out), self._stream())
        return obj

    def _phase_loss(self, c10, c12, phi12, want_loss: bool):
        torch = self.torch
        if not self._has_g:
            raise ValueError("no resident Fourier stack")
        if self._rotation is None:
            self.cache_rotation(0.0)
        phase = torch.empty((self.n, self.n), dtype=torch.float32, device=self.plan.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.plan.device) if want_loss else None
        _abi.call("ssb_phase_loss", self.plan.handle, float(c10), float(c12), float(phi12),
                  _abi.ptr(phase), _abi.ptr(loss), self._stream())
        return phase, loss

    def phase_loss_into(self, c10: float, c12: float, phi12: float, loss_out) -> None:
        """Enqueue one exact-loss evaluation writing a 1-element float32 device tensor."""
        if not self._has_g:
            raise ValueError("no resident Fourier stack")
        if loss_out is None:
            raise ValueError("loss_out is required")
        check_out(loss_out, "loss_out", 1, self.torch.float32, self.plan.device)
        if self._rotation is None:
            self.cache_rotation(0.0)
        _abi.call("ssb_phase_loss", self.plan.handle, float(c10), float(c12), float(phi12), None,
                  _abi.ptr(loss_out), self._stream())

    def _phase_loss_full(self, mags, angles, want_loss: bool):
        torch = self.torch
        if not self._has_g:
            raise ValueError("no resident Fourier stack")
        if self._rotation is None:
            self.cache_rotation(0.0)
        m = np.ascontiguousarray(np.asarray(_host(mags), dtype=np.float64).reshape(-1))
        a = np.ascontiguousarray(np.asarray(_host(angles), dtype=np.float64).reshape(-1))
        if m.size != 14 or a.size != 14:
            raise ValueError("aberration magnitudes and angles must have 14 entries")
        phase = torch.empty((self.n, self.n), dtype=torch.float32, device=self.plan.device)
        loss = torch.empty(1, dtype=torch.float32, device=self.plan.device) if want_loss else None
        _abi.call("ssb_phase_loss_full", self.plan.handle, m.ctypes.data_as(_abi._dp),
                  a.ctypes.data_as(_abi._dp), _abi.ptr(phase), _abi.ptr(loss), self._stream())
        return phase, loss

    def batch_phase_loss(self, mags, angles, rotations=None, out=None, graph: bool = False):
        """Exact losses of T trials (``mags``/``angles``: [T, 14]; ``rotations``:
        [T] rad or None) as a float32 device tensor, enqueued without a sync.
        ``graph``: replay the batch from a captured hipGraph (one upload
        synchronisation; the same bits)."""
        torch = self.torch
        if not self._has_g:
            raise ValueError("no resident Fourier stack")
        if self._rotation is None:
            self.cache_rotation(0.0)
        m = np.ascontiguousarray(np.asarray(_host(mags), dtype=np.float64).reshape(-1, 14))
        a = np.ascontiguousarray(np.asarray(_host(angles), dtype=np.float64).reshape(-1, 14))
        if m.shape != a.shape:
            raise ValueError("mags and angles must have the same [T, 14] shape")
        t = m.shape[0]
        r = None
        if rotations is not None:
            r = np.ascontiguousarray(np.asarray(_host(rotations), dtype=np.float64).reshape(-1))
            if r.size != t:
                raise ValueError("rotations must have one entry per trial")
        if out is None:
            out = torch.empty(t, dtype=torch.float32, device=self.plan.device)
        check_out(out, "out", t, torch.float32, self.plan.device)
        _abi.call("ssb_batch_phase_loss_graph" if graph else "ssb_batch_phase_loss", self.plan.handle, int(t),
                  m.ctypes.data_as(_abi._dp), a.ctypes.data_as(_abi._dp),
                  r.ctypes.data_as(_abi._dp) if r is not None else None, _abi.ptr(out), self._stream())
        return out

    def reconstruct_full(self, mags, angles):
        """Mean per-BF phase for the 14-coefficient aberration vector
        (``SSBEngine.reconstruct_full``, ``cuda/engine.py:1208-1296``)."""
        return self._phase_loss_full(mags, angles, False)[0]

    def reconstruct_full_with_loss(self, mags, angles):
        """(mean per-BF phase, loss) for the 14-coefficient vector."""
        phase, loss = self._phase_loss_full(mags, angles, True)
        return phase, float(loss.item())

    def reconstruct(self, c10: float, c12: float, phi12: float):
        """Mean of the per-BF phase images (float32 [N, N], on the GPU).

        ``SSBEngine.reconstruct`` (``cuda/engine.py:1153-1191``)."""
        return self._phase_loss(c10, c12, phi12, False)[0]

    def reconstruct_with_loss(self, c10: float, c12: float, phi12: float):
        """(mean per-BF phase, exact phase-variance loss as a Python float).

        ``SSBEngine.reconstruct_with_loss`` (``cuda/engine.py:1464-1514``);
        the loss read-back synchronises the stream, as ``float(loss)`` does."""
        phase, loss = self._phase_loss(c10, c12, phi12, True)
        return phase, float(loss.item())

    def recon_u16(self, data, c10: float, c12: float, phi12: float, rotation_rad: float = 0.0,
                  threshold: float = 0.0, bf_radius: float | None = None, obj_out=None, phase_out=None,
                  after: "HipSSBEngine | None" = None):
        """End-to-end plain SSB from resident raw frames of any supported
        sample type (no host round trip; the name keeps the uint16 fast path
        it was first written for).

        ``after``: another engine whose most recent ``after``-ordered recon's
        raw-data stream pass this one's stream pass must follow (two engines
        on two streams keep two reconstructions in flight)."""
        code = self._check_data(data)
        if after is not None and code != _DTYPE_CODES["uint16"]:
            raise ValueError("recon_u16(after=...) reads uint16 frames")
        nn = self.n * self.n
        check_out(obj_out, "obj_out", nn, self.torch.complex64, self.plan.device)
        check_out(phase_out, "phase_out", nn, self.torch.float32, self.plan.device)
        radius = float(bf_radius) if bf_radius is not None else 0.0
        if after is None:
            _abi.call("ssb_recon_typed", self.plan.handle, _abi.ptr(data), code, float(c10), float(c12),
                      float(phi12), float(rotation_rad), float(threshold), radius,
                      _abi.ptr(obj_out), _abi.ptr(phase_out), self._stream())
        else:
            _abi.call("ssb_recon_u16_after", self.plan.handle, after.plan.handle, _abi.ptr(data), float(c10),
                      float(c12), float(phi12), float(rotation_rad), float(threshold), radius,
                      _abi.ptr(obj_out), _abi.ptr(phase_out), self._stream())
        self._rotation = float(rotation_rad)
        self._has_g = False
        self._bf = None            # selected on the device; read back on demand
        self._device_bf = True

    # ------------------------------------------------------------ timing
    def tune_set(self, knob: int, value: int) -> None:
        """Set a launch-shape knob (``tuner.KNOBS``); 0 restores the default."""
        _abi.call("ssb_tune_set", self.plan.handle, int(knob), int(value))

    def tune_get(self, knob: int) -> int:
        out = ctypes.c_int(0)
        _abi.call("ssb_tune_get", self.plan.handle, int(knob), ctypes.byref(out))
        return int(out.value)

    def timing_begin(self, max_calls: int, boundaries=None) -> None:
        """Record per-stage hipEvents for the next ``max_calls`` recon calls;
        ``boundaries`` (stage indices, the stage count meaning the end) limits
        the events to those boundaries (default: all)."""
        mask = 0
        for b in boundaries or ():
            mask |= 1 << int(b)
        _abi.call("ssb_timing_begin", self.plan.handle, int(max_calls), int(mask))

    def timing_end(self) -> dict:
        """Summed per-stage milliseconds of the recorded recon calls (synchronises)."""
        lib = _abi.lib()
        n = int(lib.ssb_stage_count())
        out = np.zeros(n, dtype=np.float32)
        calls = ctypes.c_int(0)
        _abi.call("ssb_timing_end", self.plan.handle, out.ctypes.data_as(_abi._fp), n, ctypes.byref(calls))
        names = [lib.ssb_stage_name(i).decode() for i in range(n)]
        res = {name: float(v) for name, v in zip(names, out)}
        res["calls"] = int(calls.value)
        return res
