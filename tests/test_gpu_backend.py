"""The ``hip`` backend behind the reference's SSB contract, on the GPU:
construction (3D/4D input, scan padding, detector-sampling detection),
``result``/``preview``/``browser_state``, the fit, against the oracle."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import ssb_oracle as O
from tests import _cases as C
from tests._gpu import synthetic

pytestmark = pytest.mark.gpu

ABER = dict(zip(("C10", "C12", "phi12"), C.MAKE_ENGINE_ABERRATIONS))


def backend(data, **kw):
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    args = dict(voltage_kV=300, semiangle=21.4, scan_sampling=0.5, det_sampling=1.0, bf_radius=3)
    args.update(kw)
    return HipSSBBackend(data, **args)


def test_protocol_result_and_preview_match_oracle(cuda, golden):
    from quantem.gpu_amd.ssb.types import SSBProtocol
    data = C.poisson_data(123)
    ssb = backend(cuda.from_numpy(data).cuda(), aberrations=dict(ABER))
    assert isinstance(ssb, SSBProtocol)
    assert ssb.backend == "hip" and ssb.scan_shape == (128, 128) and ssb.detector_shape == (16, 16)
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    assert ssb.num_bf == p.bf.rows.size
    res = ssb.result()
    ref_obj = O.object_wave(p.G, p.dc, p.geom, *C.MAKE_ENGINE_ABERRATIONS)
    rms, p999, _ = C.phase_errors(res.phase.cpu().numpy(), np.angle(ref_obj))
    assert rms < 1e-4 and p999 < 3e-4
    _, ref_loss = O.phase_and_loss(p.G, p.dc, p.geom, *C.MAKE_ENGINE_ABERRATIONS)
    assert res.loss == pytest.approx(ref_loss, rel=1e-5, abs=1e-6)
    assert res.backend == "hip" and res.num_bf == ssb.num_bf and res.aberrations == ssb.aberrations
    phase, loss = ssb.preview(ABER, compute_loss=True, higher_order_magnitudes=None, higher_order_angles=None)
    assert phase.dtype == np.float32 and loss == pytest.approx(res.loss, rel=1e-6)
    st = ssb.browser_state()
    assert st.num_bf == ssb.num_bf and st.backend == "hip"
    np.testing.assert_allclose(st.kx_bf, p.geom.kx, atol=1e-7)
    np.testing.assert_allclose(st.aperture_k, p.geom.aperture, atol=1e-6)
    ssb.close()


def test_numpy_3d_input_and_roi96_padding(cuda, golden):
    fx = golden["fixtures"]["roi96"]
    data = C.poisson_data(fx["seed"], scan=(96, 96), add=fx["disk_add"])
    flat = data.reshape(96 * 96, 16, 16)
    ssb = backend(flat, scan_shape=(96, 96))
    assert ssb.scan_shape == (fx["padded_scan"],) * 2
    padded = O.normalize_scan(data)
    p = O.prepare(padded, bf_radius=3, **C.POISSON_KW)
    np.testing.assert_array_equal(ssb.bf_inds_row, p.bf.rows)
    np.testing.assert_array_equal(ssb.bf_inds_col, p.bf.cols)
    with pytest.raises(ValueError, match="perfect square"):
        backend(flat[:-1])


def test_detector_sampling_detection_matches_oracle(cuda):
    cfg, data = synthetic(128, 64, {"C10": 200.0})
    _, mean = O.mean_dp(data)
    (rc, cc), radius = O.detect_bf_radius(mean)
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    ssb = HipSSBBackend(cuda.from_numpy(data).cuda(), semiangle=20.0, scan_sampling=cfg.step_A,
                        det_sampling=None, voltage_kV=300)
    assert ssb.angular_sampling[0] == pytest.approx(2 * 20.0 / radius)
    yy, xx = np.mgrid[:64, :64]
    for r0 in (10, 16, 23):
        disk = np.where((yy - 30) ** 2 + (xx - 33) ** 2 <= r0 ** 2, 1000, 1).astype(np.uint16)
        d4 = np.broadcast_to(disk, (128, 128, 64, 64)).copy()
        from quantem.gpu_amd.ssb.compute.hip.engine import HipSSBEngine, PlanSpec, SSBPlan
        eng = HipSSBEngine(SSBPlan(PlanSpec(128, 64, 64, 8, 0.0197, 0.02, 1e-3, 1e-3, 0.5, 0.5)))
        eng.mean_dp(cuda.from_numpy(d4).cuda())
        assert eng.detect_bf_radius() == ((30, 33), r0)


def test_fit_recovers_defocus(cuda):
    """Synthetic weak-phase data at C10 = 200 A: search + Nelder-Mead on the
    exact objective lands near the truth and not above the truth's loss."""
    cfg, data = synthetic(128, 64, {"C10": 200.0}, dose=1e6)
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    ssb = HipSSBBackend(cuda.from_numpy(data).cuda(), semiangle=20.0, scan_sampling=cfg.step_A,
                        det_sampling=1.25, voltage_kV=300)
    res = ssb.fit(trials=64, refinement="nelder-mead",
                  search_ranges={"C10_nm": (-500.0, 500.0), "C12_nm": 0.0, "phi12_deg": 0.0},
                  refine_lock=None, seed=42, verbose=False)
    assert abs(res.aberrations["C10"] - 200.0) < 10.0
    truth = ssb.reconstruct_with_loss(200.0, 0.0, 0.0)[1]
    assert res.loss <= truth + 1e-6
    assert res.n_trials == 64 and res.refine_method == "nelder-mead" and res.refine_nfev > 4
    assert len(res.optuna_trials) == 64


def test_result_routes_higher_order_through_mean_phase(cuda):
    data = C.poisson_data(124)
    ab = {"C10": -120.0, "C12": 55.0, "phi12": math.radians(17.0), "C30": 1.5e5}
    ssb = backend(cuda.from_numpy(data).cuda(), aberrations=dict(ab))
    res = ssb.result()
    assert res.loss is None                       # cuda/backend.py:820-826
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    ref, definition = O.result_phase(p.G, p.dc, p.geom, ab)
    assert definition == "mean-bf"
    rms, p999, _ = C.phase_errors(res.phase.cpu().numpy(), ref)
    assert rms < 1e-4 and p999 < 3e-4
    np.testing.assert_allclose(res.amplitude.cpu().numpy(), 1.0, atol=1e-6)


@pytest.mark.parametrize("add", [80, 300])
def test_export_brightfield_is_exact(cuda, tmp_path, add):
    """Detector-major BF columns equal ``data.reshape(-1, k, k)[:, rows, cols].T``
    (``tests/test_ssb_cuda_128.py:291-303``); uint8 only when lossless."""
    data = C.poisson_data(123, add=add)
    ssb = backend(cuda.from_numpy(data).cuda())
    path, seconds = ssb.export_brightfield(cuda.from_numpy(data).cuda(), tmp_path / "bf")
    expect = O.gather_bf_columns(data, ssb.bf_inds_row, ssb.bf_inds_col)
    dtype = np.uint8 if data.max() <= 255 else np.uint16
    assert path.suffix == (".u8" if dtype == np.uint8 else ".u16") and seconds >= 0
    got = np.fromfile(path, dtype=dtype).reshape(ssb.num_bf, 128 * 128)
    np.testing.assert_array_equal(got, expect.astype(dtype))
    st = ssb.browser_state()
    assert st.bf_source_path == path and st.bf_source_max_value == int(data.max())


def test_grid_search_picks_the_grid_minimum(cuda):
    cfg, data = synthetic(128, 64, {"C10": 200.0}, dose=1e6)
    from quantem.gpu_amd.ssb.compute.hip.backend import HipSSBBackend
    ssb = HipSSBBackend(cuda.from_numpy(data).cuda(), semiangle=20.0, scan_sampling=cfg.step_A,
                        det_sampling=1.25, voltage_kV=300, aberrations={"C10": 180.0, "C12": 0.0, "phi12": 0.0})
    ssb.grid_search(window={"C10_nm": 40, "C12_nm": 0, "phi12_deg": 0},
                    n_points={"C10_nm": 9, "C12_nm": 1, "phi12_deg": 1}, verbose=False)
    grid = np.linspace(140.0, 220.0, 9)
    losses = [ssb.reconstruct_with_loss(c, 0.0, 0.0)[1] for c in grid]
    assert ssb.aberrations["C10"] == pytest.approx(grid[int(np.argmin(losses))])
    assert abs(ssb.aberrations["C10"] - 200.0) <= 10.0


def test_fit_save_to_reuses_an_exact_prior_result(cuda, tmp_path):
    """``fit(save_to=...)``: the second identical call loads the saved result
    (same object wave, marked reused); changed settings or ``force`` rerun."""
    data = C.poisson_data(124)
    src = tmp_path / "scan.raw"
    src.write_bytes(data.tobytes())
    ssb = backend(cuda.from_numpy(data).cuda(), aberrations=dict(ABER))
    kw = dict(trials=8, refinement=None, search_ranges={"C10_nm": (-200.0, 200.0), "C12_nm": 0.0,
                                                         "phi12_deg": 0.0},
              refine_lock=None, seed=42, verbose=False, save_to=tmp_path / "out", source_path=src)
    with pytest.raises(ValueError, match="source_path"):
        ssb.fit(**dict(kw, source_path=None))
    first = ssb.fit(**kw)
    assert not first.reused and (tmp_path / "out" / "ssb-fit.npz").is_file()
    ssb.aberrations.update(ABER)
    again = ssb.fit(**kw)
    assert again.reused
    np.testing.assert_array_equal(again.object_wave.cpu().numpy(), first.object_wave.cpu().numpy())
    assert again.aberrations == first.aberrations and ssb.aberrations["C10"] == first.aberrations["C10"]
    assert not ssb.fit(**dict(kw, seed=7)).reused
    assert not ssb.fit(**dict(kw, seed=7, force=True)).reused


def test_preview_context_uses_a_strided_bf_subset(cuda):
    """``preview_context(k)``: previews inside the block equal the FP64 oracle
    on every (full // k)-th BF pixel (the first k), with the full-disk dc;
    outside it they use the full disk again."""
    data = C.poisson_data(123)
    ssb = backend(cuda.from_numpy(data).cuda(), aberrations=dict(ABER))
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    full = p.bf.rows.size
    k = full // 3
    idx = np.arange(0, full, max(1, full // k))[:k]
    ctx = ssb.preview_context(k)
    assert ctx.num_bf == k
    with ctx:
        phase, loss = ssb.preview(ABER, compute_loss=True, higher_order_magnitudes=None,
                                  higher_order_angles=None)
    geom = O.detector_geometry(p.bf.rows[idx], p.bf.cols[idx], p.bf.center, (16, 16), (1.0, 1.0),
                               p.geom.wavelength, 21.4, 0.0, 128, 0.5)
    ref_phase, ref_loss = O.phase_and_loss(p.G[idx], p.dc, geom, *C.MAKE_ENGINE_ABERRATIONS)
    rms, p999, _ = C.phase_errors(phase, ref_phase)
    assert rms < 1e-4 and p999 < 3e-4
    assert loss == pytest.approx(ref_loss, rel=1e-4, abs=1e-6)
    full_phase, _ = ssb.preview(ABER, compute_loss=False, higher_order_magnitudes=None, higher_order_angles=None)
    ref_full, _ = O.phase_and_loss(p.G, p.dc, p.geom, *C.MAKE_ENGINE_ABERRATIONS)
    assert C.phase_errors(full_phase, ref_full)[0] < 1e-4


def test_bf_subsample_searches_on_the_subset(cuda):
    data = C.poisson_data(125)
    ssb = backend(cuda.from_numpy(data).cuda(), aberrations=dict(ABER))
    ranges = {"C10_nm": (-150.0, 150.0), "C12_nm": 0.0, "phi12_deg": 0.0}
    ssb.optimize(aberrations=ranges, n_trials=8, verbose=False, bf_subsample=0.5)
    hist = ssb._optuna_trials
    assert len(hist) == 8 and ssb._subset_engine is None
    with ssb._bf_subset(np.arange(0, ssb.num_bf, 2)):
        again = ssb._evaluate_params([dict(h["params"], C12_nm=0.0, phi12_deg=0.0, C30=0.0,
                                           rotation_angle_deg=0.0) for h in hist[:3]])
    np.testing.assert_allclose(again, [h["loss"] for h in hist[:3]], rtol=1e-6)
    with pytest.raises(ValueError, match="bf_subsample"):
        ssb.optimize(aberrations=ranges, n_trials=4, verbose=False, bf_subsample=-1.0)


def test_search_over_uneven_datasets_matches_single_searches(cuda):
    """Scans of 128 (native), 96 (padded to 128) and 192 (padded to 256)
    searched together give each dataset the losses of its own search."""
    from quantem.gpu_amd.ssb.compute.hip import multi
    datasets = [C.poisson_data(123), C.poisson_data(124, scan=(96, 96)), C.poisson_data(125, scan=(192, 192))]
    bks = [backend(cuda.from_numpy(d).cuda(), aberrations=dict(ABER)) for d in datasets]
    assert [b.scan_shape[0] for b in bks] == [128, 128, 256]
    ranges = {"C10_nm": (-150.0, 150.0), "C12_nm": 0.0, "phi12_deg": 0.0, "C30": 0.0,
              "rotation_angle_deg": (-2.0, 2.0)}
    res = multi.search_datasets(bks, ranges, n_trials=12, batch=5)
    for b, r in zip(bks, res):
        single = b._evaluate_params([dict(h["params"], C12_nm=0.0, phi12_deg=0.0, C30=0.0)
                                     for h in r["history"]])
        np.testing.assert_array_equal(single, [h["loss"] for h in r["history"]])
        assert r["loss"] == min(h["loss"] for h in r["history"])
