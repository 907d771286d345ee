"""Build libquantem_ssb_hip.so for gfx950 with hipcc (in-tree, no JIT cache).

Usage: ``python -m quantem.gpu_amd.csrc.build`` or ``build()`` from Python.
The library lands in ``quantem/gpu_amd/lib/``.  Compilation of the
translation units runs in parallel; the link step combines them.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
REPO = PKG.parent.parent
LIB_DIR = PKG / "lib"
LIB_NAME = "libquantem_ssb_hip.so"
SOURCES = ["ssb_prepare.hip", "ssb_object.hip", "ssb_loss.hip", "ssb_learn.hip", "ssb_bslz4.hip", "ssb_bin.hip", "ssb_api.hip",
           "ssb_trials.hip", "ssb_noalloc.hip"]

ARCH = os.environ.get("QSSB_OFFLOAD_ARCH", "gfx950")

COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are disabled for every HIP translation unit
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    f"-I{REPO / 'include'}",
    f"-I{HERE}",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found; ROCm is required to build the SSB library")
    return exe


def build(verbose: bool = False) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    hipcc = _hipcc()
    build_dir = PKG / "_build"
    build_dir.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    sources = [HERE / s for s in SOURCES if (HERE / s).exists()]

    def compile_one(src: Path) -> Path:
        obj = build_dir / (src.stem + ".o")
        cmd = [hipcc, *COMMON_FLAGS, "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stdout}\n{res.stderr}")
        return obj

    with ThreadPoolExecutor(max_workers=min(len(sources), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, sources))
    out = LIB_DIR / LIB_NAME
    tmp = LIB_DIR / (LIB_NAME + ".tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
