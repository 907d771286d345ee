"""Sample types other than uint16, and the detector gain, through the C ABI
against the oracle.

The reference keeps the caller's dtype (``cuda/backend.py:195-201``), sums
integer frames in uint64 and float frames in float64 (``:430-441``), rounds
every BF sample to float32 before the scan FFT (``:570-571``) and applies
``detector_gain`` to the mean DP and to the BF samples (``:220-226``,
``442-443``, ``574-576``).  Integer sums, masks and padding must be
bit-exact; float32 frames are summed in double-float here (the mean within
one float32 ulp of the float64 sum's); phases within 1e-4 rad RMS of FP64.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ssb_oracle as O
from tests import _cases as C

pytestmark = pytest.mark.gpu

PHASE_RMS = 1e-4


def _typed(base: np.ndarray, kind: str) -> np.ndarray:
    """The Poisson fixture in another sample type, with values that need it."""
    b = base.astype(np.int64)
    if kind == "uint8":
        return np.minimum(b, 255).astype(np.uint8)
    if kind == "int8":
        return np.minimum(b, 127).astype(np.int8)
    if kind == "int16":
        return (b * 250).astype(np.int16)                       # beyond int8, inside int16
    if kind == "uint32":
        return (b * 70001 + 3).astype(np.uint32)                # beyond uint16
    if kind == "uint32-wide":
        return (b * 4_000_037 + 11).astype(np.uint32)           # beyond float32's 2^24: samples round
    if kind == "int32":
        return (b * 90001 - 5).astype(np.int32)
    if kind == "float32":
        return (b.astype(np.float32) * np.float32(0.7316) + np.float32(0.125)).astype(np.float32)
    raise ValueError(kind)


KINDS = ["uint8", "int8", "int16", "uint32", "uint32-wide", "int32", "float32"]


def _engine(cuda, n=128, det=16):
    from tests._gpu import make_engine
    return make_engine(cuda, n, (det, det), 64, O.wavelength_angstrom(300), 21.4, (1.0, 1.0), 0.5)


@pytest.mark.parametrize("kind", KINDS)
def test_mean_dp_and_mask_of_other_sample_types(cuda, kind):
    data = _typed(C.poisson_data(123), kind)
    eng = _engine(cuda)
    d = cuda.from_numpy(data).cuda()
    sums, mean = eng.mean_dp(d)
    rs, rm = O.mean_dp(data)
    if kind == "float32":
        assert sums is None
        got = mean.cpu().numpy()
        assert np.all(np.abs(got.view(np.int32).astype(np.int64) - rm.view(np.int32).astype(np.int64)) <= 1)
    else:
        np.testing.assert_array_equal(sums.cpu().numpy().astype(np.uint64), rs)
        np.testing.assert_array_equal(mean.cpu().numpy().view(np.uint32), rm.view(np.uint32))
    ref = O.select_bf(rm, 0.0, 3.0)
    bf = eng.select_bf(0.0, 3.0)
    np.testing.assert_array_equal(bf.rows, ref.rows)
    np.testing.assert_array_equal(bf.cols, ref.cols)


def test_signed_sums_wrap_like_the_reference(cuda):
    """Negative samples enter the uint64 sum sign-extended (modulo 2^64)."""
    rng = np.random.default_rng(7)
    data = rng.integers(-3000, 3000, size=(128, 128, 8, 8), dtype=np.int16)
    eng = _engine(cuda, det=8)
    sums, mean = eng.mean_dp(cuda.from_numpy(data).cuda())
    rs, rm = O.mean_dp(data)
    np.testing.assert_array_equal(sums.cpu().numpy().astype(np.uint64), rs)
    np.testing.assert_array_equal(mean.cpu().numpy().view(np.uint32), rm.view(np.uint32))


@pytest.mark.parametrize("kind", KINDS)
def test_reconstruction_of_other_sample_types(cuda, kind):
    data = _typed(C.poisson_data(124), kind)
    c10, c12, phi12 = C.MAKE_ENGINE_ABERRATIONS
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    ref = np.angle(O.object_wave(p.G, p.dc, p.geom, c10, c12, phi12))
    eng = _engine(cuda)
    d = cuda.from_numpy(data).cuda()
    phase = cuda.empty((128, 128), dtype=cuda.float32, device="cuda")
    eng.recon_u16(d, c10, c12, phi12, 0.0, 0.0, 3.0, phase_out=phase)
    rms, p999, _ = C.phase_errors(phase.cpu().numpy(), ref)
    assert rms < PHASE_RMS and p999 < 3e-4, (rms, p999)
    assert eng.num_bf == p.bf.rows.size
    assert eng.bf_state().dc == pytest.approx(p.dc, rel=1e-6)
    # the staged path: resident stack, then the object
    eng.mean_dp(d)
    eng.select_bf(0.0, 3.0)
    eng.prepare_g(d)
    eng.cache_rotation(0.0)
    staged = cuda.empty_like(phase)
    eng.reconstruct_object(c10, c12, phi12, phase_out=staged)
    assert C.phase_errors(staged.cpu().numpy(), ref)[0] < PHASE_RMS


@pytest.mark.parametrize("kind", ["uint8", "int16", "uint32", "int32", "float32"])
def test_scan_padding_of_other_sample_types(cuda, kind):
    from quantem.gpu_amd.ssb.compute.hip.engine import normalize_scan
    data = _typed(C.poisson_data(125, scan=(96, 100)), kind)
    got = normalize_scan(cuda.from_numpy(data).cuda()).cpu().numpy()
    ref = O.normalize_scan(data)
    assert got.dtype == ref.dtype and got.shape == ref.shape == (128, 128, 16, 16)
    if kind == "float32":
        inside = got[16:112, 14:114]
        np.testing.assert_array_equal(inside, data)
        d = np.abs(got.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
        assert d.max() <= 1
    else:
        np.testing.assert_array_equal(got, ref)


def test_binned_uint32_frames_reconstruct(cuda):
    """``bin_frames``' default uint32 output goes straight into the backend
    (detector 32 binned to 16), matching the oracle on the binned frames."""
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    from quantem.gpu_amd.ssb.compute.hip.binning import bin_frames
    raw = C.poisson_data(126, det=32, radius=8, add=20)
    binned = bin_frames(cuda.from_numpy(raw.reshape(-1, 32, 32)).cuda(), 2)
    assert binned.dtype == cuda.uint32
    host = binned.cpu().numpy().reshape(128, 128, 16, 16)
    np.testing.assert_array_equal(host, O.bin_sum(raw, 2).astype(np.uint32))
    ssb = HipSSBBackend(binned.reshape(128, 128, 16, 16), semiangle=21.4, scan_sampling=0.5, det_sampling=2.0,
                        voltage_kV=300, bf_radius=3.0, aberrations={"C10": -120.0})
    res = ssb.result(compute_loss=False)
    p = O.prepare(host, voltage_kV=300, semiangle_mrad=21.4, scan_sampling=0.5, det_sampling=2.0, bf_radius=3)
    ref = np.angle(O.object_wave(p.G, p.dc, p.geom, -120.0, 0.0, 0.0))
    assert C.phase_errors(res.phase.cpu().numpy(), ref)[0] < PHASE_RMS
    ssb.close()


@pytest.mark.parametrize("kind", ["uint16", "float32"])
def test_detector_gain(cuda, kind):
    """A per-pixel gain changes the mean DP (mask) and the BF samples exactly
    as the reference applies it; the result matches the oracle with the same
    gain and differs from the ungained one."""
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    base = C.poisson_data(127)
    data = base if kind == "uint16" else _typed(base, "float32")
    gain = np.random.default_rng(11).uniform(0.6, 1.4, size=(16, 16)).astype(np.float32)
    ab = {"C10": -120.0, "C12": 55.0, "phi12": 0.3}
    ssb = HipSSBBackend(cuda.from_numpy(data).cuda(), semiangle=21.4, scan_sampling=0.5, det_sampling=1.0,
                        voltage_kV=300, bf_radius=3.0, aberrations=dict(ab), detector_gain=gain)
    _, mean = ssb._engine.mean_dp(cuda.from_numpy(data).cuda())
    _, rm = O.mean_dp(data, gain)
    if kind == "uint16":
        np.testing.assert_array_equal(mean.cpu().numpy().view(np.uint32), rm.view(np.uint32))
    p = O.prepare(data, bf_radius=3, detector_gain=gain, **C.POISSON_KW)
    np.testing.assert_array_equal(ssb.bf_inds_row, p.bf.rows)
    np.testing.assert_array_equal(ssb.bf_inds_col, p.bf.cols)
    assert ssb.dc_value == pytest.approx(p.dc, rel=1e-6)
    res = ssb.result(compute_loss=True)
    ref = np.angle(O.object_wave(p.G, p.dc, p.geom, ab["C10"], ab["C12"], ab["phi12"]))
    assert C.phase_errors(res.phase.cpu().numpy(), ref)[0] < PHASE_RMS
    _, ref_loss = O.phase_and_loss(p.G, p.dc, p.geom, ab["C10"], ab["C12"], ab["phi12"])
    assert res.loss == pytest.approx(ref_loss, rel=1e-4, abs=1e-6)
    p0 = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    plain = np.angle(O.object_wave(p0.G, p0.dc, p0.geom, ab["C10"], ab["C12"], ab["phi12"]))
    assert C.phase_errors(plain, ref)[0] > 10 * PHASE_RMS
    with pytest.raises(ValueError, match="detector_gain shape"):
        HipSSBBackend(cuda.from_numpy(data).cuda(), semiangle=21.4, scan_sampling=0.5, det_sampling=1.0,
                      voltage_kV=300, detector_gain=np.ones((8, 8), np.float32))
    ssb.close()


def test_uint8_brightfield_export(cuda, tmp_path):
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    data = _typed(C.poisson_data(123), "uint8")
    ssb = HipSSBBackend(cuda.from_numpy(data).cuda(), semiangle=21.4, scan_sampling=0.5, det_sampling=1.0,
                        voltage_kV=300, bf_radius=3.0)
    path, _ = ssb.export_brightfield(data, tmp_path / "bf")
    assert path.suffix == ".u8"
    got = np.fromfile(path, dtype=np.uint8).reshape(ssb.num_bf, -1)
    np.testing.assert_array_equal(got, O.gather_bf_columns(data, ssb.bf_inds_row, ssb.bf_inds_col))
    with pytest.raises(TypeError):
        ssb.export_brightfield(data.astype(np.int16), tmp_path / "bad")
    ssb.close()


def test_unsupported_sample_type_raises(cuda):
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    with pytest.raises(NotImplementedError, match="uint8, uint16"):
        HipSSBBackend(np.zeros((128, 128, 16, 16), np.int64), semiangle=21.4, scan_sampling=0.5,
                      det_sampling=1.0, voltage_kV=300)
