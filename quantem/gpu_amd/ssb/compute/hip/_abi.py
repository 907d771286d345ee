"""ctypes binding of ``libquantem_ssb_hip.so`` (declarations: ``include/quantem_ssb.h``).

The library is loaded from the package's ``lib/`` directory (built in-tree by
``quantem.gpu_amd.csrc.build``).  There is no fallback: if the library or a
GPU is missing, loading raises.  Status codes map to the exception types of
the reference CUDA backend (``cuda/kernels/common.py:783-796``,
``cuda/backend.py:808-816``, ``cuda/engine.py:1276-1280``).
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_NAME = "libquantem_ssb_hip.so"
LIB_PATH = Path(__file__).resolve().parents[3] / "lib" / LIB_NAME

QSSB_OK = 0
QSSB_ERR_ARG = -1
QSSB_ERR_OOM = -2
QSSB_ERR_UNSUPPORTED = -3
QSSB_ERR_LAUNCH = -4
QSSB_ERR_STATE = -5

_c_int = ctypes.c_int
_c_float = ctypes.c_float
_c_double = ctypes.c_double
_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_ip = ctypes.POINTER(ctypes.c_int32)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_pp = ctypes.POINTER(ctypes.c_void_p)

# name -> argtypes; every function returns int
SIGNATURES: dict[str, list] = {
    "ssb_version": [],
    "ssb_plan_create": [_c_int, _c_int, _c_int, _c_int, _c_double, _c_double, _c_double, _c_double,
                        _c_double, _c_double, _pp],
    "ssb_plan_destroy": [_vp],
    "ssb_set_detector_sampling": [_vp, _c_double, _c_double],
    "ssb_set_scan_frequencies": [_vp, _fp, _fp],
    "ssb_mean_dp_u16": [_vp, _vp, _i64, _vp, _vp, _vp],
    "ssb_select_bf": [_vp, _c_float, _c_float, _vp],
    "ssb_set_bf": [_vp, _ip, _ip, _c_int, _c_float, _c_float, _c_float, _c_float, _vp],
    "ssb_get_bf": [_vp, _ip, _ip, _c_int, ctypes.POINTER(_c_int), _fp, _vp],
    "ssb_cache_rotation": [_vp, _c_double, _vp],
    "ssb_get_geometry": [_vp, _vp, _vp, _vp, _c_int, _vp],
    "ssb_prepare_g_u16": [_vp, _vp, _vp],
    "ssb_export_g": [_vp, _vp, _c_int, _vp],
    "ssb_import_g": [_vp, _vp, _c_int, _c_int, _vp],
    "ssb_object": [_vp, _c_double, _c_double, _c_double, _vp, _vp, _vp],
    "ssb_recon_u16": [_vp, _vp, _c_double, _c_double, _c_double, _c_double, _c_float, _c_float, _vp, _vp,
                      _vp],
    "ssb_recon_u16_after": [_vp, _vp, _vp, _c_double, _c_double, _c_double, _c_double, _c_float, _c_float, _vp,
                            _vp, _vp],
    "ssb_stage_count": [],
    "ssb_bslz4_workspace_bytes": [_c_int, ctypes.c_longlong, _c_int],
    "ssb_bslz4_encode_bound": [ctypes.c_longlong, _c_int],
    "ssb_bslz4_encode_u16": [_vp, ctypes.c_longlong, _c_int, _vp, ctypes.c_longlong],
    "ssb_bslz4_decode_u16": [_vp, ctypes.c_longlong, _vp, _vp, _c_int, ctypes.c_longlong, _c_int, _vp, _vp,
                             ctypes.c_longlong, ctypes.POINTER(_c_int), _vp],
    "ssb_bin_u16": [_vp, ctypes.c_longlong, _c_int, _c_int, _c_int, _vp, _c_int, _vp, _vp],
    "ssb_tune_candidates": [_c_int, _c_int, ctypes.POINTER(_c_int), _c_int],
    "ssb_tune_set": [_vp, _c_int, _c_int],
    "ssb_tune_get": [_vp, _c_int, ctypes.POINTER(_c_int)],
    "ssb_timing_begin": [_vp, _c_int, _c_int],
    "ssb_timing_end": [_vp, _fp, _c_int, ctypes.POINTER(_c_int)],
    "ssb_phase_loss": [_vp, _c_double, _c_double, _c_double, _vp, _vp, _vp],
    "ssb_phase_loss_full": [_vp, _dp, _dp, _vp, _vp, _vp],
    "ssb_batch_phase_loss": [_vp, _c_int, _dp, _dp, _dp, _vp, _vp],
    "ssb_scan_target": [_c_int, _c_int],
    "ssb_learn_features": [_vp, _ip, _ip, _vp, _vp],
    "ssb_max_u16": [_vp, _i64, ctypes.POINTER(ctypes.c_uint), _vp],
    "ssb_bf_columns_u16": [_vp, _vp, _c_int, _c_int, _c_int, _vp, _vp],
    "ssb_detect_bf_radius": [_vp, _c_double, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                             ctypes.POINTER(_c_int), _vp],
    "ssb_normalize_scan_u16": [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp],
    "ssb_dtype_size": [_c_int],
    "ssb_mean_dp_typed": [_vp, _vp, _c_int, _i64, _vp, _vp, _vp],
    "ssb_set_detector_gain": [_vp, _vp, _vp],
    "ssb_prepare_g_typed": [_vp, _vp, _c_int, _vp],
    "ssb_recon_typed": [_vp, _vp, _c_int, _c_double, _c_double, _c_double, _c_double, _c_float, _c_float, _vp,
                        _vp, _vp],
    "ssb_normalize_scan_typed": [_vp, _c_int, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp],
    "ssb_bf_columns_u8": [_vp, _vp, _c_int, _c_int, _vp, _vp],
    "ssb_normalize_scan_u16_noalloc": [_vp, _c_int, _c_int, _c_int, _c_int, _vp, _c_int, _vp],
    "ssb_batch_phase_loss_graph": [_vp, _c_int, _dp, _dp, _dp, _vp, _vp],
    "ssb_trials_create": [_pp, _c_int, _c_int, _pp],
    "ssb_trials_destroy": [_vp],
    "ssb_trials_set": [_vp, _c_int, _dp, _dp, _dp, _vp],
    "ssb_trials_phase_variance": [_vp, _vp, _c_int, _vp],
}

_lib: ctypes.CDLL | None = None


class SSBLibraryError(RuntimeError):
    """The native SSB library is missing or cannot be loaded."""


def lib() -> ctypes.CDLL:
    """Load (once) and return the native library."""
    global _lib
    if _lib is not None:
        return _lib
    path = Path(os.environ.get("QSSB_LIBRARY", LIB_PATH))
    if not path.exists():
        raise SSBLibraryError(
            f"{path} not found. Build it with `python -m quantem.gpu_amd.csrc.build`; "
            "the hip SSB backend has no CPU fallback."
        )
    handle = ctypes.CDLL(str(path))
    for name, argtypes in SIGNATURES.items():
        fn = getattr(handle, name, None)
        if fn is None:
            continue
        fn.argtypes = argtypes
        fn.restype = _c_int
    handle.ssb_last_error.argtypes = []
    handle.ssb_last_error.restype = ctypes.c_char_p
    handle.ssb_stage_name.argtypes = [_c_int]
    handle.ssb_stage_name.restype = ctypes.c_char_p
    for name in ("ssb_bslz4_workspace_bytes", "ssb_bslz4_encode_bound", "ssb_bslz4_encode_u16"):
        if hasattr(handle, name):
            getattr(handle, name).restype = ctypes.c_longlong
    _lib = handle
    return handle


def last_error() -> str:
    msg = lib().ssb_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> None:
    """Raise the reference's exception type for a non-zero status."""
    if rc == QSSB_OK:
        return
    msg = last_error() or what
    if rc == QSSB_ERR_ARG or rc == QSSB_ERR_STATE:
        raise ValueError(msg)
    if rc == QSSB_ERR_OOM:
        raise MemoryError(
            f"Out of GPU memory in the hip SSB backend ({msg}). Try a smaller bf_radius "
            "or release other GPU allocations."
        )
    if rc == QSSB_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RuntimeError(msg or f"hip SSB call failed with status {rc}")


def call(name: str, *args) -> None:
    fn = getattr(lib(), name)
    check(fn(*args), name)


def ptr(tensor) -> ctypes.c_void_p:
    """Device pointer of a torch tensor (or None)."""
    if tensor is None:
        return ctypes.c_void_p(0)
    return ctypes.c_void_p(tensor.data_ptr())


def stream_handle(stream=None) -> ctypes.c_void_p:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)
