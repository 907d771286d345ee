"""Host-side behaviour that needs no GPU: no CPU fallback, the one-GPU
bench contract, the entry points and the synthetic data contract."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_engine_refuses_to_run_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from quantem.gpu_amd.ssb.compute.hip.engine import PlanSpec, SSBPlan
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        SSBPlan(PlanSpec(128, 16, 16, 8, 0.0197, 0.02, 1e-3, 1e-3, 0.5, 0.5))


def test_product_path_does_not_import_the_oracle():
    pkg = ROOT / "quantem" / "gpu_amd"
    for path in pkg.rglob("*.py"):
        text = path.read_text()
        assert "import oracle" not in text and "from oracle" not in text, path


def test_bench_refuses_more_than_one_gpu():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 2
    assert "one MI355X" in res.stderr


def test_graft_entry_points_exist():
    sys.path.insert(0, str(ROOT))
    import __graft_entry__ as g
    assert callable(g.build) and callable(g.smoke)


def test_synthetic_dataset_contract():
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate, object_phase
    cfg = SyntheticConfig(scan=16, det=32, dose=1e4, seed=3, aberrations={"C10": 200.0})
    a = generate(cfg, "cpu")
    b = generate(cfg, "cpu")
    assert a.dtype == np.uint16 and a.shape == (16, 16, 32, 32)
    np.testing.assert_array_equal(a, b)                       # seeded
    per_frame = a.reshape(256, -1).sum(axis=1)
    assert abs(per_frame.mean() / 1e4 - 1.0) < 0.02           # dose per probe
    ph = object_phase(cfg)
    assert abs(np.sqrt(np.mean(ph ** 2)) - 0.1) < 1e-12       # 0.1 rad RMS
    assert cfg.wavelength == pytest.approx(0.019687489006849, rel=1e-12)


def test_objective_modes_follow_the_reference_contract():
    """``cuda/backend.py:1021-1026``: unknown modes are a ValueError; the
    sparse-row "native" objective is not built and says so."""
    from quantem.gpu_amd.ssb.compute.hip.backend import _check_objective_mode
    _check_objective_mode("exact")
    with pytest.raises(NotImplementedError, match="native"):
        _check_objective_mode("native")
    with pytest.raises(ValueError, match="objective_mode"):
        _check_objective_mode("fast")


def test_bin_frames_checks_its_input_before_any_gpu_call():
    torch = pytest.importorskip("torch")
    from quantem.gpu_amd.ssb.compute.hip.binning import bin_frames
    with pytest.raises(TypeError):
        bin_frames(torch.zeros((2, 8, 8), dtype=torch.uint16), 2)            # host tensor
    with pytest.raises(TypeError):
        bin_frames(np.zeros((2, 8, 8), dtype=np.uint16), 2)


def test_output_checks_reject_host_short_and_mistyped_tensors():
    """Caller-supplied outputs are checked before any native call: a host
    tensor, a wrong dtype, a short or non-contiguous tensor raise ValueError."""
    torch = pytest.importorskip("torch")
    from quantem.gpu_amd.ssb.compute.hip.engine import check_out
    dev = "cuda:0"
    check_out(None, "out", 4, torch.float32, dev)
    with pytest.raises(ValueError, match="must live on"):
        check_out(torch.zeros(4), "out", 4, torch.float32, dev)
    with pytest.raises(ValueError, match="dtype"):
        check_out(torch.zeros(4, dtype=torch.float64), "out", 4, torch.float32, dev)
    with pytest.raises(ValueError, match="torch tensor"):
        check_out(np.zeros(4, np.float32), "out", 4, torch.float32, dev)
