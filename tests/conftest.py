"""Shared pytest configuration.

``gpu`` tests need a ROCm GPU and the in-tree HIP library
(``quantem/gpu_amd/lib/libquantem_ssb_hip.so``); they compare the HIP path,
called through the C ABI, with the CPU oracle.  Everything else runs on CPU.
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN = ROOT / "tests" / "golden" / "reference_vectors.json"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def golden() -> dict:
    return json.loads(GOLDEN.read_text())


@pytest.fixture(scope="session")
def cuda():
    """The torch module, with a GPU present (skips on CPU-only hosts)."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU on this host")
    torch.cuda.set_device(0)
    return torch
