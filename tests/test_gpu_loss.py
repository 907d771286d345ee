"""GPU parity of the exact per-BF phase path (mean of per-BF phases and the
phase-variance loss) against the FP64 oracle, through the C ABI.

Bars follow the reference's own tests of this path
(``tests/test_ssb_cuda_128.py:336-419``): at 128² ``assert_allclose`` with
rtol = atol = 2e-4 on the phase and rel 1e-5 / abs 1e-6 on the loss
(``tests/ssb_precision.py:7-10``); at larger sizes, where a pixel may cross
the atan2 branch cut differently and shift the mean by 2*pi/N_bf, the phase
p99.9 < 3e-4 and the loss within rel 1e-4 / abs 1e-4.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import ssb_oracle as O
from tests import _cases as C
from tests._gpu import engine_for_case, engine_for_data, synthetic

pytestmark = pytest.mark.gpu

ABER = C.MAKE_ENGINE_ABERRATIONS


def phase_loss(eng, c10, c12, phi12):
    phase, loss = eng.reconstruct_with_loss(c10, c12, phi12)
    return phase.cpu().numpy(), loss


def test_make_engine_128_matches_reference_bars(cuda):
    case = C.make_engine_case(128, 7)
    phase, loss = phase_loss(engine_for_case(cuda, case), *ABER)
    ref_phase, ref_loss = O.phase_and_loss(case.G, case.dc, case.geometry(), *ABER)
    np.testing.assert_allclose(phase, ref_phase, rtol=2e-4, atol=2e-4)
    assert loss == pytest.approx(ref_loss, rel=1e-5, abs=1e-6)


@pytest.mark.parametrize("size,num_bf", [(256, 5), (512, 5), (1024, 3)])
def test_make_engine_large_matches_reference_bars(cuda, size, num_bf):
    case = C.make_engine_case(size, num_bf)
    phase, loss = phase_loss(engine_for_case(cuda, case), *ABER)
    ref_phase, ref_loss = O.phase_and_loss(case.G, case.dc, case.geometry(), *ABER)
    assert np.percentile(np.abs(phase - ref_phase), 99.9) < 3e-4
    assert loss == pytest.approx(ref_loss, rel=1e-4, abs=1e-4)


@pytest.mark.parametrize("size,num_bf", [(128, 7), (256, 5), (512, 3)])
def test_hermitian_and_full_storage_agree(cuda, size, num_bf):
    full = C.hermitian_stack(size, num_bf)
    half = np.ascontiguousarray(full[:, :, : size // 2 + 1])
    out = []
    for g in (O.expand_hermitian(half), half):
        case = C.make_engine_case(size, num_bf, g_qk=g)
        out.append(phase_loss(engine_for_case(cuda, case), *ABER))
    ref_phase, ref_loss = O.phase_and_loss(half, case.dc, case.geometry(), *ABER)
    for phase, loss in out:
        assert np.percentile(np.abs(phase - ref_phase), 99.9) < 3e-4
        assert loss == pytest.approx(ref_loss, rel=1e-4, abs=1e-4)
    assert np.percentile(np.abs(out[0][0] - out[1][0]), 99.9) < 3e-4


def test_zero_input_known_answer(cuda, golden):
    za = golden["zero_input"]
    case = C.make_engine_case(256, 1, g_qk=np.zeros((1, 256, 129), np.complex64))
    case.dc = complex(*za["dc"])
    phase, loss = phase_loss(engine_for_case(cuda, case), 0.0, 0.0, 0.0)
    assert np.abs(phase).max() == 0.0
    assert loss == pytest.approx(za["loss"], abs=za["loss_atol"])


@pytest.mark.parametrize("seed", [123, 125])
def test_poisson_fixture_phase_and_loss(cuda, seed, golden):
    fx = golden["fixtures"]["poisson"]
    data = C.poisson_data(seed)
    p = O.prepare(data, bf_radius=fx["bf_radius"], **C.POISSON_KW)
    ref_phase, ref_loss = O.phase_and_loss(p.G, p.dc, p.geom, *ABER)
    eng = engine_for_data(cuda, data, semiangle_mrad=21.4, det_sampling=1.0, step=0.5)
    d = cuda.from_numpy(data).cuda()
    eng.mean_dp(d)
    eng.select_bf(0.0, fx["bf_radius"])
    eng.prepare_g(d)
    eng.cache_rotation(0.0)
    phase, loss = phase_loss(eng, *ABER)
    rms, p999, _ = C.phase_errors(phase, ref_phase)
    assert rms < 1e-4 and p999 < 3e-4
    assert loss == pytest.approx(ref_loss, rel=1e-5, abs=1e-6)
    # mean phase alone equals the phase of the combined call; repeat is bit-identical
    again, loss2 = phase_loss(eng, *ABER)
    np.testing.assert_array_equal(again, phase)
    assert loss2 == loss
    np.testing.assert_array_equal(eng.reconstruct(*ABER).cpu().numpy(), phase)


def test_synthetic_loss_is_smallest_near_true_defocus(cuda):
    """Weak-phase data at C10 = 200 A (scan 128, detector 64): the exact loss
    matches FP64 at three defoci and is lowest at the true one."""
    cfg, data = synthetic(128, 64, {"C10": 200.0})
    p = O.prepare(data, voltage_kV=300, semiangle_mrad=20.0, scan_sampling=cfg.step_A,
                  det_sampling=1.25)
    eng = engine_for_data(cuda, data, semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    d = cuda.from_numpy(data).cuda()
    eng.mean_dp(d)
    eng.select_bf(0.0, None)
    eng.prepare_g(d)
    eng.cache_rotation(0.0)
    losses = {}
    for c10 in (0.0, 200.0, 400.0):
        phase, loss = phase_loss(eng, c10, 0.0, 0.0)
        ref_phase, ref_loss = O.phase_and_loss(p.G, p.dc, p.geom, c10, 0.0, 0.0)
        rms, p999, _ = C.phase_errors(phase, ref_phase)
        assert rms < 1e-4 and p999 < 3e-4
        assert loss == pytest.approx(ref_loss, rel=1e-5, abs=1e-6)
        losses[c10] = loss
    assert losses[200.0] < losses[0.0] and losses[200.0] < losses[400.0]


# --------------------------------------------------------- full (C_s) chi
def _vec(**kw):
    mags, angs, _ = O.aberration_vectors(kw)
    return mags, angs


def test_full_chi_reduces_to_quadratic_path(cuda):
    case = C.make_engine_case(128, 7)
    eng = engine_for_case(cuda, case)
    c10, c12, phi12 = ABER
    mags, angs = _vec(C10=c10, C12=c12, phi12=phi12)
    p_full, l_full = eng.reconstruct_full_with_loss(mags, angs)
    p_quad, l_quad = eng.reconstruct_with_loss(c10, c12, phi12)
    assert np.percentile(np.abs(p_full.cpu().numpy() - p_quad.cpu().numpy()), 99.9) < 1e-5
    assert l_full == pytest.approx(l_quad, rel=1e-5, abs=1e-7)


@pytest.mark.parametrize("ab", [
    {"C10": -120.0, "C30": 2.0e5},
    {"C10": 60.0, "C12": 30.0, "phi12": 0.3, "C21_mag": 4.0e3, "C21_angle": 25.0, "C23_mag": 2.0e3,
     "C23_angle": -40.0, "C30": -1.0e5},
])
def test_full_chi_matches_oracle(cuda, ab):
    data = C.poisson_data(123)
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    mags, angs = _vec(**ab)
    ref_phase, ref_loss = O.phase_and_loss_full(p.G, p.dc, p.geom, mags, angs)
    eng = engine_for_data(cuda, data, semiangle_mrad=21.4, det_sampling=1.0, step=0.5)
    d = cuda.from_numpy(data).cuda()
    eng.mean_dp(d)
    eng.select_bf(0.0, 3.0)
    eng.prepare_g(d)
    eng.cache_rotation(0.0)
    phase, loss = eng.reconstruct_full_with_loss(mags, angs)
    rms, p999, _ = C.phase_errors(phase.cpu().numpy(), ref_phase)
    assert rms < 1e-4 and p999 < 3e-4
    assert loss == pytest.approx(ref_loss, rel=1e-4, abs=1e-5)


def test_full_chi_loss_is_smallest_near_true_cs(cuda):
    """Weak-phase data with C_s = 1e5 A (10 um) and C10 = 100 A."""
    cfg, data = synthetic(128, 64, {"C10": 100.0, "C30": 1.0e5}, dose=1e6)
    eng = engine_for_data(cuda, data, semiangle_mrad=20.0, det_sampling=1.25, step=cfg.step_A)
    d = cuda.from_numpy(data).cuda()
    eng.mean_dp(d)
    eng.select_bf(0.0, None)
    eng.prepare_g(d)
    eng.cache_rotation(0.0)
    p = O.prepare(data, voltage_kV=300, semiangle_mrad=20.0, scan_sampling=cfg.step_A, det_sampling=1.25)
    losses = {}
    for cs in (-1.0e5, 1.0e5, 3.0e5):
        mags, angs = _vec(C10=100.0, C30=cs)
        phase, loss = eng.reconstruct_full_with_loss(mags, angs)
        ref_phase, ref_loss = O.phase_and_loss_full(p.G, p.dc, p.geom, mags, angs)
        rms, p999, _ = C.phase_errors(phase.cpu().numpy(), ref_phase)
        assert rms < 1e-4 and p999 < 3e-4
        assert loss == pytest.approx(ref_loss, rel=1e-4, abs=1e-5)
        losses[cs] = loss
    assert losses[1.0e5] < losses[-1.0e5] and losses[1.0e5] < losses[3.0e5]


def test_batch_of_trials_equals_single_evaluations(cuda):
    """One batched call (quadratic, C_s and rotated trials) gives the same
    bits as the trials evaluated one at a time."""
    data = C.poisson_data(125)
    eng = engine_for_data(cuda, data, semiangle_mrad=21.4, det_sampling=1.0, step=0.5)
    d = cuda.from_numpy(data).cuda()
    eng.mean_dp(d)
    eng.select_bf(0.0, 3.0)
    eng.prepare_g(d)
    eng.cache_rotation(0.0)
    trials = [(dict(C10=-120.0, C12=55.0, phi12=0.3), 0.0), (dict(C10=40.0), 0.0),
              (dict(C10=40.0, C30=1.0e5), 0.0), (dict(C10=-60.0), math.radians(4.0)),
              (dict(C10=-60.0, C30=-5.0e4), math.radians(-3.0))]
    mags = np.stack([_vec(**t)[0] for t, _ in trials])
    angs = np.stack([_vec(**t)[1] for t, _ in trials])
    rots = np.array([r for _, r in trials])
    batch = eng.batch_phase_loss(mags, angs, rots).cpu().numpy()
    single = []
    for (t, r), m, a in zip(trials, mags, angs):
        eng.cache_rotation(r)
        if np.any(m[2:] != 0):
            single.append(eng.reconstruct_full_with_loss(m, a)[1])
        else:
            single.append(eng.reconstruct_with_loss(m[0], m[1], a[1])[1])
    eng.cache_rotation(0.0)
    np.testing.assert_array_equal(batch, np.asarray(single, dtype=np.float32))
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    for (t, r), m, a, got in zip(trials, mags, angs, batch):
        geom = O.detector_geometry(p.bf.rows, p.bf.cols, p.bf.center, (16, 16), (1.0, 1.0), p.geom.wavelength,
                                   21.4, r, 128, 0.5)
        ref = O.phase_and_loss_full(p.G, p.dc, geom, m, a)[1]
        assert got == pytest.approx(ref, rel=1e-4, abs=1e-5)
