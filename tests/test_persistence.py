"""Saved-result persistence (the storage contract of
``src/quantem/gpu/ssb/_persistence.py``), on the CPU."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from quantem.gpu_amd.ssb import persistence as P
from quantem.gpu_amd.ssb.types import SSBResult


def _result(seed=0):
    rng = np.random.default_rng(seed)
    wave = (rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))).astype(np.complex64)
    return SSBResult(object_wave=wave, backend="hip", aberrations={"C10": 12.5, "C12": 0.0, "phi12": 0.1},
                     rotation_angle_deg=2.0, loss=0.0123, n_trials=16, num_bf=97,
                     scan_sampling_A=0.25, bf_center=(63.5, 64.0), bf_radius=16.0)


def test_result_paths_follow_the_reference_rules(tmp_path):
    assert P.result_paths(tmp_path / "out", "fit") == (tmp_path / "out" / "ssb-fit.npz",
                                                        tmp_path / "out" / "ssb-fit.json")
    assert P.result_paths(tmp_path / "a.npz", "reconstruct") == (tmp_path / "a.npz", tmp_path / "a.json")
    with pytest.raises(ValueError, match="directory or an .npz file"):
        P.result_paths(tmp_path / "a.h5", "fit")


def test_json_value_is_deterministic_and_strict():
    v = P.json_value({"b": np.float32(1.5), "a": (1, np.arange(2)), "p": Path("/x")})
    assert list(v) == ["a", "b", "p"] and v["a"] == [1, [0, 1]] and v["b"] == 1.5 and v["p"] == "/x"
    with pytest.raises(TypeError, match="Cannot save"):
        P.json_value({"x": object()})


def test_save_then_load_round_trips_exactly(tmp_path):
    paths = P.result_paths(tmp_path, "fit")
    sig = {"operation": "fit", "settings": {"trials": 16}}
    res = P.save_result(_result(), paths=paths, signature=sig, input_metadata={"source_path": "x"})
    assert res.saved_path == paths[0] and not res.reused
    meta = json.loads(paths[1].read_text())
    assert meta["schema"] == P.SCHEMA and meta["result"]["loss"] == pytest.approx(0.0123)
    back = P.load_result(paths=paths, signature=sig)
    assert back is not None and back.reused
    np.testing.assert_array_equal(back.object_wave, _result().object_wave)
    assert back.aberrations == _result().aberrations and back.bf_center == (63.5, 64.0)
    assert back.n_trials == 16 and back.num_bf == 97


def test_any_mismatch_returns_none(tmp_path):
    paths = P.result_paths(tmp_path, "fit")
    sig = {"settings": {"trials": 16}}
    P.save_result(_result(), paths=paths, signature=sig, input_metadata={})
    assert P.load_result(paths=paths, signature={"settings": {"trials": 17}}) is None
    meta = json.loads(paths[1].read_text())
    meta["schema"] = P.SCHEMA + 1
    paths[1].write_text(json.dumps(meta))
    assert P.load_result(paths=paths, signature=sig) is None
    paths[1].write_text("{not json")
    assert P.load_result(paths=paths, signature=sig) is None
    assert P.load_result(paths=P.result_paths(tmp_path / "none", "fit"), signature=sig) is None


def test_reuse_or_run_runs_once_unless_forced(tmp_path):
    calls = []

    def run():
        calls.append(1)
        return _result(len(calls))

    sig = {"settings": 1}
    a = P.reuse_or_run("fit", tmp_path, sig, run, input_metadata={})
    b = P.reuse_or_run("fit", tmp_path, sig, run, input_metadata={})
    assert len(calls) == 1 and b.reused
    np.testing.assert_array_equal(a.object_wave, b.object_wave)
    c = P.reuse_or_run("fit", tmp_path, sig, run, input_metadata={}, force=True)
    assert len(calls) == 2 and not c.reused
    d = P.reuse_or_run("fit", tmp_path, {"settings": 2}, run, input_metadata={})
    assert len(calls) == 3 and not d.reused


def test_path_signature_describes_without_reading(tmp_path):
    f = tmp_path / "scan.h5"
    f.write_bytes(b"x" * 10)
    sig = P.path_signature(f)
    assert sig["files"][0]["size"] == 10
    assert P.path_signature(tmp_path / "missing")["missing"] is True
    s1 = P.software_signature()
    assert s1 == P.software_signature() and len(s1["source_sha256"]) == 64


def test_publish_leaves_only_the_target(tmp_path):
    store = P.ResultStore.at(tmp_path / "deep" / "dir", "reconstruct")
    store.save(_result(3), {"k": 1}, {})
    store.save(_result(4), {"k": 1}, {})                     # replaced in place
    names = sorted(p.name for p in (tmp_path / "deep" / "dir").iterdir())
    assert names == ["ssb-reconstruct.json", "ssb-reconstruct.npz"]
    back = store.load({"k": 1})
    np.testing.assert_array_equal(back.object_wave, _result(4).object_wave)


def test_signature_comparison_is_canonical(tmp_path):
    # key order, tuples vs lists and numpy scalars do not change a signature
    store = P.ResultStore.at(tmp_path, "fit")
    store.save(_result(), {"b": (1, 2), "a": np.float64(0.5)}, {})
    assert store.load({"a": 0.5, "b": [1, 2]}) is not None
    assert store.load({"a": 0.5, "b": [1, 3]}) is None


def test_git_state_and_directory_signature(tmp_path):
    state = P._git_state()
    if state is not None:
        assert len(state["commit"]) == 40 and isinstance(state["dirty"], bool)
    (tmp_path / "b").mkdir()
    (tmp_path / "b" / "z.bin").write_bytes(b"1")
    (tmp_path / "a.bin").write_bytes(b"22")
    sig = P.path_signature(tmp_path)
    paths = [f["path"] for f in sig["files"]]
    assert paths[0] == str(tmp_path) and paths[1:] == sorted(paths[1:]) and len(paths) == 3
