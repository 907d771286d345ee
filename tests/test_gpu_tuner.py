"""Launch-shape knobs on the GPU: every candidate of every knob gives the
defaults' phase (bit for bit for the integer stream pass; to float rounding
where the FFT and trotter code is instantiated for another tile shape), and
the autotuner's cache round-trips."""
from __future__ import annotations

import pytest

from tests._gpu import engine_for_data, synthetic
from quantem.gpu_amd.ssb.compute.hip import tuner

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("scan,det", [(128, 64), (256, 32)])
def test_every_candidate_gives_the_same_phase(cuda, scan, det):
    cfg, data = synthetic(scan, det, {"C10": 150.0})
    eng = engine_for_data(cuda, data, semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    d = cuda.from_numpy(data).cuda()
    ref = cuda.empty((scan, scan), dtype=cuda.float32, device="cuda")
    eng.recon_u16(d, 150.0, 0.0, 0.0, phase_out=ref)
    out = cuda.empty_like(ref)
    for knob, code in tuner.KNOBS.items():
        for v in tuner.candidates(scan, knob):
            eng.tune_set(code, v)
            assert eng.tune_get(code) == v
            out.fill_(0.0)
            eng.recon_u16(d, 150.0, 0.0, 0.0, phase_out=out)
            if knob == "stream_wg":
                assert cuda.equal(out, ref), (knob, v)
            else:
                # another instantiation may contract a multiply-add differently
                assert float((out - ref).abs().max()) < 2e-6, (knob, v)
        eng.tune_set(code, 0)
    with pytest.raises(ValueError, match="candidate"):
        eng.tune_set(tuner.KNOBS["gather_pairs"], 3)


def test_autotune_sets_and_caches_a_choice(cuda, tmp_path):
    cfg, data = synthetic(128, 64, {"C10": 150.0})
    eng = engine_for_data(cuda, data, semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    d = cuda.from_numpy(data).cuda()
    path = tmp_path / "tuning.json"
    res = tuner.autotune(eng, d, c10=150.0, repeats=2, store=True, path=path)
    for knob, v in res["config"].items():
        assert v in tuner.candidates(128, knob)
        assert eng.tune_get(tuner.KNOBS[knob]) == v
        assert len(res["ms"][knob]) == len(tuner.candidates(128, knob))
    for code in tuner.KNOBS.values():
        eng.tune_set(code, 0)
    got = tuner.load_cached(eng, eng.num_bf, path=path)
    assert got == res["config"]
    assert all(eng.tune_get(tuner.KNOBS[k]) == v for k, v in got.items())
