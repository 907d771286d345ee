"""The C ABI: the in-tree library loads and exports every symbol that
``include/*.h`` declares; argument checks that run before any device call
map to the reference's exception types.  No compute runs here."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from quantem.gpu_amd.ssb.compute.hip import _abi

ROOT = Path(__file__).resolve().parents[1]


_RET = {"int": ctypes.c_int, "long long": ctypes.c_longlong, "const char*": ctypes.c_char_p}


def declared() -> dict[str, object]:
    """Every entry point the headers declare, with the ctypes type of its return value."""
    out = {}
    for header in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        for ret, name in re.findall(r"^\s*(int|long long|const char\s*\*)\s+((?:ssb|qnn)_\w+)\s*\(", text,
                                    flags=re.M):
            assert name not in out, f"{name} declared twice"
            out[name] = _RET[re.sub(r"\s*\*", "*", re.sub(r"\s+", " ", ret.strip()))]
    return out


def declared_symbols() -> list[str]:
    return list(declared())


@pytest.fixture(scope="module")
def lib():
    if not _abi.LIB_PATH.exists():
        pytest.fail(f"{_abi.LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`")
    return _abi.lib()


def test_header_declares_the_boundary():
    names = declared_symbols()
    for core in ("ssb_plan_create", "ssb_plan_destroy", "ssb_mean_dp_u16", "ssb_select_bf",
                 "ssb_prepare_g_u16", "ssb_object", "ssb_recon_u16", "ssb_import_g", "ssb_export_g"):
        assert core in names
    assert len(names) == len(set(names))


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared_symbols() if getattr(lib, n, None) is None]
    assert not missing, f"not exported: {missing}"
    decl = declared()
    for n in _abi.SIGNATURES:
        if n in decl:
            assert getattr(lib, n).restype is decl[n], n


def test_version_and_stage_names(lib):
    assert lib.ssb_version() >= 1
    n = lib.ssb_stage_count()
    names = [lib.ssb_stage_name(i).decode() for i in range(n)]
    assert names[:2] == ["superset", "mean_dp_stream"] and "colfft_trotter" in names


def test_argument_errors_map_to_reference_exceptions(lib):
    plan = ctypes.c_void_p()
    with pytest.raises(NotImplementedError, match="scan size 64"):
        _abi.call("ssb_plan_create", 64, 64, 64, 100, 0.0197, 0.02, 1e-3, 1e-3, 0.25, 0.25,
                  ctypes.byref(plan))
    with pytest.raises(ValueError, match="positive"):
        _abi.call("ssb_plan_create", 128, 64, 64, 100, -1.0, 0.02, 1e-3, 1e-3, 0.25, 0.25,
                  ctypes.byref(plan))
    with pytest.raises(NotImplementedError, match="65536"):
        _abi.call("ssb_plan_create", 128, 512, 512, 100, 0.0197, 0.02, 1e-3, 1e-3, 0.25, 0.25,
                  ctypes.byref(plan))
    with pytest.raises(ValueError):
        _abi.call("ssb_plan_create", 128, 64, 64, 100, 0.0197, 0.02, 1e-3, 1e-3, 0.25, 0.25, None)
    assert plan.value is None
    with pytest.raises(ValueError):
        _abi.call("ssb_object", None, 0.0, 0.0, 0.0, None, None, None)


def test_status_mapping():
    _abi.check(_abi.QSSB_OK)
    with pytest.raises(MemoryError, match="Out of GPU memory"):
        _abi.check(_abi.QSSB_ERR_OOM, "x")
    with pytest.raises(RuntimeError):
        _abi.check(_abi.QSSB_ERR_LAUNCH, "x")


def test_scan_target_matches_reference_rule(lib):
    # cuda/backend.py:238-246: next of 128/256/512/1024 at or above the longest side
    cases = {(96, 96): 128, (128, 128): 128, (129, 5): 256, (130, 40): 256, (300, 512): 512,
             (513, 1): 1024, (1100, 10): 1024, (0, 4): 0}
    for (h, w), t in cases.items():
        assert lib.ssb_scan_target(h, w) == t
