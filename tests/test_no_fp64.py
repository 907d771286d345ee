"""No FP64 arithmetic on the GPU: every HIP translation unit of the library,
compiled for gfx950 with the library's own flags, contains no float64 VALU
instruction (the arithmetic is FP32 with compensated sums and double-float
pairs; host code may use double freely)."""
from __future__ import annotations

import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from quantem.gpu_amd.csrc import build as B

F64_OP = re.compile(r"^\s+v_\w*_f64\w*", re.MULTILINE)


def _device_asm(src) -> str:
    cmd = [B._hipcc(), *[f for f in B.COMMON_FLAGS if f != "-fPIC"], "--cuda-device-only", "-S", str(src), "-o", "-"]
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stderr[-2000:]
    return res.stdout


@pytest.mark.skipif(shutil.which("hipcc") is None and not B.Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc is not installed")
def test_device_code_has_no_float64_instructions():
    sources = [B.HERE / s for s in B.SOURCES]
    with ThreadPoolExecutor(max_workers=8) as pool:
        asm = dict(zip(sources, pool.map(_device_asm, sources)))
    found = {src.name: sorted(set(m.group(0).strip() for m in F64_OP.finditer(text)))[:5]
             for src, text in asm.items() if F64_OP.search(text)}
    assert not found, f"float64 instructions in device code: {found}"
    # the scan finds device code where the kernels are (some units hold host code only)
    assert all("v_" in asm[B.HERE / s] for s in ("ssb_prepare.hip", "ssb_object.hip", "ssb_trials.hip"))


def test_the_scan_sees_float64_instructions():
    # the pattern the scan uses, on lines of the assembler's own format
    text = "\tv_add_f32_e32 v1, v2, v3\n\tv_fma_f64 v[0:1], v[2:3], v[4:5], v[6:7]\n\tv_cvt_f64_f32_e32 v[8:9], v1\n"
    assert sorted(m.group(0).strip() for m in F64_OP.finditer(text)) == ["v_cvt_f64_f32_e32", "v_fma_f64"]
    assert not F64_OP.search("\tv_add_f32_e32 v1, v2, v3\n\ts_add_u32 s0, s1, 64\n")
