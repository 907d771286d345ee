"""Pins of the CPU oracle (``oracle/ssb_oracle.py``) against the reference's
test vectors, closed forms, brute force on tiny inputs and invariants.

Each check is chosen so a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand, a missing conjugate) fails it.
"""
from __future__ import annotations

import cmath
import math

import numpy as np
import pytest

from oracle import ssb_oracle as O
from tests import _cases as C


# ----------------------------------------------------------------- optics
def test_wavelength_table_and_relativistic_formula(golden):
    tab = golden["wavelength"]["table_A"]
    for kv, lam in tab.items():
        assert O.wavelength_angstrom(float(kv)) == lam
    k = golden["wavelength"]["codata"]
    for kv, lam in tab.items():
        v = float(kv) * 1e3
        rel = k["h"] / math.sqrt(2 * k["m_e"] * k["e"] * v * (1 + k["e"] * v / (2 * k["m_e"] * k["c"] ** 2)))
        assert rel * 1e10 == pytest.approx(lam, rel=1e-9)
    # off-table voltages use the formula and stay monotone between table entries
    assert tab["300"] > O.wavelength_angstrom(300.5) > tab["400"]


def test_scan_frequencies_closed_form():
    q = O.scan_frequencies(8, 0.5)
    np.testing.assert_array_equal(q, np.array([0, 1, 2, 3, -4, -3, -2, -1]) / 4.0)
    assert O.scan_frequencies(7, 2.0)[-1] == -1.0 / 14.0


# --------------------------------------------------------- integer stages
def test_mean_dp_sums_exact_above_float32_limit():
    rng = np.random.default_rng(7)
    data = rng.integers(0, 65536, size=(300, 1, 5, 6), dtype=np.uint16)
    data[:, :, 0, 0] = 65535                      # 65535*300 > 2^24
    sums, mean = O.mean_dp(data)
    ref = [[sum(int(v) for v in data[:, 0, i, j]) for j in range(6)] for i in range(5)]
    assert sums.dtype == np.uint64
    assert sums.tolist() == ref
    assert int(sums[0, 0]) == 65535 * 300
    const = np.full((4, 5, 3, 3), 1234, dtype=np.uint16)
    np.testing.assert_array_equal(O.mean_dp(const)[1], np.full((3, 3), 1234.0, np.float32))


def test_normalize_scan_pads_roi96_to_128(golden):
    fx = golden["fixtures"]["roi96"]
    data = C.poisson_data(fx["seed"], scan=tuple(fx["shape"][:2]), add=fx["disk_add"])
    out = O.normalize_scan(data)
    assert out.shape == (128, 128, 16, 16) and out.dtype == np.uint16
    np.testing.assert_array_equal(out[16:112, 16:112], data)
    sums = data.reshape(-1, 256).astype(np.int64).sum(axis=0)
    floor_mean = (sums // (96 * 96)).reshape(16, 16).astype(np.uint16)
    for frame in (out[0, 0], out[127, 127], out[15, 50], out[112, 3]):
        np.testing.assert_array_equal(frame, floor_mean)
    assert O.normalize_scan(out) is out
    odd = C.poisson_data(3, scan=(130, 40), det=4, radius=1)
    o2 = O.normalize_scan(odd)
    assert o2.shape[:2] == (256, 256)
    np.testing.assert_array_equal(o2[63:193, 108:148], odd)


def test_normalize_scan_center_crops_oversize_axis():
    data = (np.arange(1100 * 10, dtype=np.int64).reshape(1100, 10, 1, 1) % 65536).astype(np.uint16)
    out = O.normalize_scan(data)
    assert out.shape == (1024, 1024, 1, 1)
    np.testing.assert_array_equal(out[:, 507:517], data[38:38 + 1024])


def _disk_mean(shape, center, radius, inside=100.0, outside=1.0):
    yy, xx = np.mgrid[: shape[0], : shape[1]]
    disk = (yy - center[0]) ** 2 + (xx - center[1]) ** 2 <= radius ** 2
    return np.where(disk, inside, outside).astype(np.float32), disk


def test_select_bf_auto_radius_hand_built_disk():
    mean, disk = _disk_mean((32, 32), (12, 9), 5)
    bf = O.select_bf(mean)
    count = int(disk.sum())
    assert bf.center == (12.0, 9.0)
    assert bf.radius == pytest.approx(math.sqrt(count / math.pi))
    keep = [(r, c) for r in range(32) for c in range(32)
            if (r - 12) ** 2 + (c - 9) ** 2 <= count / math.pi]
    assert list(zip(bf.rows.tolist(), bf.cols.tolist())) == keep     # raster order


def test_select_bf_threshold_and_weighted_centroid():
    mean, _ = _disk_mean((20, 20), (10, 10), 3)
    mean[10, 13] = 400.0                          # pulls the weighted centroid right
    bf = O.select_bf(mean, threshold=0.2, bf_radius=10.0)
    w = np.where(mean > 80.0, mean, 0.0)
    rr, cc = np.nonzero(w)
    exp_c = (rr * w[rr, cc]).sum() / w.sum(), (cc * w[rr, cc]).sum() / w.sum()
    assert bf.center == pytest.approx(exp_c, rel=1e-6)
    assert bf.rows.size == rr.size                # every candidate within radius 10
    with pytest.raises(ValueError, match="No bright-field pixels found"):
        O.select_bf(np.zeros((8, 8), np.float32))
    with pytest.raises(ValueError, match="No bright-field pixels within"):
        O.select_bf(mean, threshold=0.2, bf_radius=1e-3)


# ---------------------------------------------------------- Fourier stack
def test_extract_gqk_reference_vector(golden):
    gv = golden["extract_gqk_arange"]
    data = np.arange(8 * 8 * 6 * 6, dtype=np.uint16).reshape(8, 8, 6, 6)
    g, dc = O.extract_gqk(data, np.array(gv["rows"]), np.array(gv["cols"]))
    expect = np.array(gv["g_half_real"]) + 1j * np.array(gv["g_half_imag"])
    assert g.shape == tuple(gv["g_half_shape"]) == (4, 8, 5)
    np.testing.assert_allclose(g, expect, rtol=0, atol=1e-9 * np.abs(expect).max())
    assert dc == pytest.approx(complex(*gv["dc"]), rel=1e-14)
    full = O.expand_hermitian(g)
    np.testing.assert_allclose(full[:, :, :5], g)


def test_extract_gqk_brute_force_dft():
    rng = np.random.default_rng(0)
    data = rng.integers(0, 1000, size=(6, 6, 3, 3)).astype(np.uint16)
    rows, cols = np.array([0, 1, 2]), np.array([2, 1, 0])
    g, _ = O.extract_gqk(data, rows, cols)
    for b in range(3):
        img = data[:, :, rows[b], cols[b]].astype(np.float64)
        for u in range(6):
            for v in range(4):
                ref = sum(img[r, c] * cmath.exp(-2j * math.pi * (u * r + v * c) / 6)
                          for r in range(6) for c in range(6))
                assert abs(g[b, u, v] - ref) < 1e-9


def test_expand_hermitian_equals_full_fft():
    rng = np.random.default_rng(1)
    real = rng.standard_normal((3, 16, 16))
    full = np.fft.fft2(real)
    np.testing.assert_allclose(O.expand_hermitian(full[:, :, :9]), full, atol=1e-12)


def test_gather_bf_columns_and_bin_sum_exact():
    rng = np.random.default_rng(2)
    data = rng.integers(0, 65536, size=(2, 3, 10, 11), dtype=np.uint16)
    rows, cols = np.array([1, 4, 9]), np.array([0, 10, 5])
    cols_out = O.gather_bf_columns(data, rows, cols)
    assert cols_out.dtype == np.uint16 and cols_out.shape == (3, 6)
    for b in range(3):
        assert cols_out[b].tolist() == data[:, :, rows[b], cols[b]].reshape(-1).tolist()
    binned = O.bin_sum(data, 3)
    assert binned.shape == (2, 3, 3, 3) and binned.dtype == np.uint64
    for i in range(3):
        for j in range(3):
            ref = data[..., 3 * i:3 * i + 3, 3 * j:3 * j + 3].astype(np.int64).sum(axis=(-2, -1))
            np.testing.assert_array_equal(binned[..., i, j], ref)


def test_bin_sum_masks_dead_pixels_and_casts_like_the_loaders():
    """``_bin_sum`` (``qg/io/backends/cpu/reference.py:27-42``, mask at
    ``:96-111``) by brute force: dead pixels count 0, edges trimmed; the
    uint16 cast keeps the low 16 bits (four 65535 counts -> 65532), the
    uint32 output is the exact sum."""
    rng = np.random.default_rng(5)
    data = rng.integers(0, 65536, size=(2, 2, 7, 9), dtype=np.uint16)
    data[0, 0, :2, :2] = 65535
    mask = rng.random((7, 9)) < 0.2
    mask[:2, :2] = False
    for factor in (1, 2, 3):
        exact = O.bin_sum(data, factor, mask)
        kr, kc = 7 // factor, 9 // factor
        assert exact.shape == (2, 2, kr, kc) and exact.dtype == np.uint64
        for i in range(kr):
            for j in range(kc):
                blk = data[..., i * factor:(i + 1) * factor, j * factor:(j + 1) * factor].astype(np.int64)
                keep = ~mask[i * factor:(i + 1) * factor, j * factor:(j + 1) * factor]
                np.testing.assert_array_equal(exact[..., i, j], (blk * keep).sum(axis=(-2, -1)))
        np.testing.assert_array_equal(O.bin_sum(data, factor, mask, np.uint32), exact.astype(np.uint32))
        np.testing.assert_array_equal(O.bin_sum(data, factor, mask, np.uint16), (exact % 65536).astype(np.uint16))
    assert int(O.bin_sum(data, 2, None, np.uint16)[0, 0, 0, 0]) == (4 * 65535) % 65536 == 65532


# ------------------------------------------------------ geometry and chi
def test_detector_rotation_is_a_rotation_of_k():
    case = C.make_engine_case(128, 7)
    g0 = case.geometry(0.0)
    th = math.radians(23.0)
    g1 = case.geometry(th)
    z0 = g0.kx + 1j * g0.ky
    z1 = g1.kx + 1j * g1.ky
    np.testing.assert_allclose(z1, z0 * cmath.exp(1j * th), atol=1e-14)
    np.testing.assert_allclose(g1.aperture, g0.aperture, atol=1e-12)   # isotropic sampling
    np.testing.assert_allclose(g0.kx, (case.rows - 15.5) / 32.0)


def test_chi_closed_forms():
    rng = np.random.default_rng(3)
    lam = 0.0197
    dx, dy = rng.uniform(-2, 2, 50), rng.uniform(-2, 2, 50)
    k2 = dx * dx + dy * dy
    a2 = k2 * lam * lam
    c2 = (dx * dx - dy * dy) / k2
    s2 = 2 * dx * dy / k2
    np.testing.assert_allclose(O.chi_c10c12(a2, c2, s2, 200.0, 0.0, 0.0, lam, np.float64),
                               math.pi * lam * k2 * 200.0, rtol=1e-13)
    # two independent statements of the C10/C12 chi agree
    mags = np.zeros(14)
    angs = np.zeros(14)
    mags[0], mags[1], angs[1] = -120.0, 55.0, math.radians(17.0)
    np.testing.assert_allclose(O.chi_full(dx, dy, lam, mags, angs),
                               O.chi_c10c12(a2, c2, s2, -120.0, 55.0, math.radians(17.0), lam,
                                            np.float64), rtol=1e-12)
    # textbook spherical aberration: (pi/2) lambda^3 |k|^4 C30
    mags = np.zeros(14)
    mags[4] = 1e5
    np.testing.assert_allclose(O.chi_full(dx, dy, lam, mags, np.zeros(14)),
                               0.5 * math.pi * lam ** 3 * k2 * k2 * 1e5, rtol=1e-12)


def test_aberration_vectors_packing():
    mags, angs, active = O.aberration_vectors({"C10": 10.0, "C12": 2.0, "phi12": 0.3})
    assert (mags[0], mags[1], angs[1], active) == (10.0, 2.0, 0.3, False)
    mags, angs, active = O.aberration_vectors({"C30": 5.0, "C23_mag": 3.0, "C23_angle": 90.0})
    assert active and mags[4] == 5.0 and mags[3] == 3.0 and angs[3] == pytest.approx(math.pi / 2)


def _symmetric_geometry():
    """BF pixels in +-k pairs about the centre (15.5, 15.5)."""
    rows = np.array([13, 18, 14, 17, 15, 16], dtype=np.int32)
    cols = np.array([14, 17, 18, 13, 15, 16], dtype=np.int32)
    case = C.make_engine_case(128, 6)
    case.rows, case.cols = rows, cols
    return case.geometry(0.0)


def test_gamma_pair_symmetry():
    g = _symmetric_geometry()
    gh = O.gamma_hat_block(g, 0, 6, -120.0, 55.0, math.radians(17.0))
    for a, b in ((0, 1), (2, 3), (4, 5)):
        # guarded points (|Gamma|^2 <= 1e-16, e.g. Q = 0) hold rounding noise * 1e8
        np.testing.assert_allclose(gh[b], -np.conj(gh[a]), atol=1e-8)
        big = np.abs(gh[a]) > 1e-6
        np.testing.assert_allclose(gh[b][big], -np.conj(gh[a][big]), atol=1e-12)


def test_gamma_zero_aberration_is_real_unit():
    g = _symmetric_geometry()
    gh = O.gamma_hat_block(g, 0, 6, 0.0, 0.0, 0.0)
    assert np.abs(gh.imag).max() < 1e-12
    big = np.abs(gh.real) > 1e-6
    np.testing.assert_allclose(np.abs(gh.real[big]), 1.0, atol=1e-12)
    assert big.mean() > 0.05


# --------------------------------------------------------------- objects
def _scalar_object(G, dc, rows, cols, center, n, recip, lam, semi_mrad, ang_mrad, q_step,
                   c10, c12, phi12):
    """Plain-loop restatement of the corrected Fourier sum and its inverse DFT."""
    qf = [(k if k < (n + 1) // 2 else k - n) / (n * q_step) for k in range(n)]
    semi = semi_mrad * 1e-3
    ay, ax = ang_mrad[0] * 1e-3, ang_mrad[1] * 1e-3

    def probe(vx, vy):
        r = math.hypot(vx, vy)
        alpha = lam * r
        if r > 1e-15:
            den = math.hypot(vx * ay, vy * ax) / r
            phi = math.atan2(vy, vx)
        else:
            den, phi = 0.0, 0.0
        a = 1.0 if den <= 1e-15 else min(1.0, max(0.0, (semi - alpha) / den + 0.5))
        chi = (math.pi / lam) * alpha * alpha * (c10 + c12 * math.cos(2.0 * (phi - phi12)))
        return a * cmath.exp(-1j * chi)

    nb = len(rows)
    S = [[0j] * n for _ in range(n)]
    for b in range(nb):
        kx = (rows[b] - center[0]) * recip[0]
        ky = (cols[b] - center[1]) * recip[1]
        pk = probe(kx, ky)
        for i in range(n):
            for j in range(n):
                qx, qy = qf[i], qf[j]
                gam = probe(qx - kx, qy - ky) * pk.conjugate() - probe(qx + kx, qy + ky).conjugate() * pk
                m2 = gam.real * gam.real + gam.imag * gam.imag
                gam = gam / math.sqrt(m2) if m2 > 1e-16 else gam * 1e8
                S[i][j] += G[b, i, j] * gam.conjugate()
    for i in range(n):
        for j in range(n):
            S[i][j] /= nb
    S[0][0] = dc
    obj = np.zeros((n, n), complex)
    for x in range(n):
        for y in range(n):
            obj[x, y] = sum(S[i][j] * cmath.exp(2j * math.pi * (i * x + j * y) / n)
                            for i in range(n) for j in range(n)) / (n * n)
    return obj


def test_object_matches_plain_loop_restatement():
    n = 8
    rng = np.random.default_rng(11)
    G = (rng.standard_normal((3, n, n)) + 1j * rng.standard_normal((3, n, n)))
    rows = np.array([14, 15, 17], dtype=np.int32)
    cols = np.array([16, 13, 15], dtype=np.int32)
    case = C.make_engine_case(n, 3, g_qk=G)
    case.rows, case.cols, case.q_step = rows, cols, 2.0
    g = case.geometry(0.0)
    c10, c12, phi12 = 3000.0, 900.0, 0.4
    obj = O.object_wave(G, case.dc, g, c10, c12, phi12)
    ref = _scalar_object(G, case.dc, rows, cols, case.center, n, case.recip, case.wavelength,
                         case.semiangle_mrad, case.ang_mrad, case.q_step, c10, c12, phi12)
    np.testing.assert_allclose(obj, ref, atol=1e-12)


@pytest.mark.parametrize("dtype,abs_tol,rel_tol", [(np.float64, 1e-15, 1e-10),
                                                    (np.float32, 5e-9, 1e-4)])
def test_fourier_sum_equals_mean_of_per_bf_ifft(dtype, abs_tol, rel_tol, golden):
    case = C.make_engine_case(128, 7)
    g = case.geometry(0.0, dtype)
    c10, c12, phi12 = C.MAKE_ENGINE_ABERRATIONS
    obj = O.object_wave(case.G, case.dc, g, c10, c12, phi12).astype(np.complex128)
    gh = O.gamma_hat_block(g, 0, 7, c10, c12, phi12)
    corr = case.G.astype(gh.dtype) * np.conj(gh)
    corr[:, 0, 0] = case.dc
    ref = np.fft.ifft2(corr.astype(np.complex128)).mean(axis=0)
    err = np.abs(obj - ref)
    rel = err / np.maximum(np.abs(ref), 1e-6)
    assert np.percentile(err, 99.9) < abs_tol
    assert np.percentile(rel, 99.9) < rel_tol


def test_zero_input_known_answer(golden):
    za = golden["zero_input"]
    n = 128
    case = C.make_engine_case(n, 1, g_qk=np.zeros((1, n, n // 2 + 1), np.complex64))
    case.dc = complex(*za["dc"])
    g = case.geometry(0.0)
    s = O.corrected_fourier_mean(case.G, case.dc, g, 0.0, 0.0, 0.0)
    assert s[0, 0] == 1.0 and np.count_nonzero(s) == 1
    obj = O.object_wave(case.G, case.dc, g, 0.0, 0.0, 0.0)
    np.testing.assert_allclose(obj, np.full((n, n), 1.0 / n ** 2), atol=2e-7 / 1e3)
    phase, loss = O.phase_and_loss(case.G, case.dc, g, 0.0, 0.0, 0.0)
    assert loss == pytest.approx(za["loss"], abs=za["loss_atol"])
    assert np.abs(phase).max() == 0.0


def test_phase_and_loss_moments():
    case = C.make_engine_case(128, 3)
    g = case.geometry(0.0)
    c10, c12, phi12 = C.MAKE_ENGINE_ABERRATIONS
    phase, loss = O.phase_and_loss(case.G, case.dc, g, c10, c12, phi12)
    gh = O.gamma_hat_block(g, 0, 3, c10, c12, phi12)
    corr = case.G.astype(np.complex128) * np.conj(gh)
    corr[:, 0, 0] = case.dc
    ph = np.angle(np.fft.ifft2(corr))
    np.testing.assert_allclose(phase, ph.mean(axis=0), atol=1e-12)
    assert loss == pytest.approx(float(ph.var(axis=0).mean()), rel=1e-10)
    one = C.make_engine_case(128, 1)
    assert O.phase_and_loss(one.G, one.dc, one.geometry(), c10, c12, phi12)[1] == pytest.approx(0, abs=1e-15)


def test_hermitian_and_full_storage_agree(golden):
    full = C.hermitian_stack(128, 5)
    half = np.ascontiguousarray(full[:, :, :65])
    case_f = C.make_engine_case(128, 5, g_qk=full)
    case_h = C.make_engine_case(128, 5, g_qk=half)
    c10, c12, phi12 = C.MAKE_ENGINE_ABERRATIONS
    of = O.object_wave(case_f.G, case_f.dc, case_f.geometry(), c10, c12, phi12)
    oh = O.object_wave(case_h.G, case_h.dc, case_h.geometry(), c10, c12, phi12)
    np.testing.assert_allclose(oh, of, atol=1e-14)


def test_fp32_restatement_tracks_fp64_on_poisson_fixture(golden):
    fx = golden["fixtures"]["poisson"]
    data = C.poisson_data(fx["seeds"][0])
    c10, c12, phi12 = C.MAKE_ENGINE_ABERRATIONS
    out = {}
    for dt in (np.float64, np.float32):
        p = O.prepare(data, dtype=dt, bf_radius=fx["bf_radius"], **C.POISSON_KW)
        out[dt] = np.angle(O.object_wave(p.G, p.dc, p.geom, c10, c12, phi12))
    rms, p999, _ = C.phase_errors(out[np.float32], out[np.float64])
    assert p999 < fx["phase_abs_p999"]
    assert rms < 1e-4


def test_object_of_weak_phase_tracks_ground_truth():
    """A weak-phase object: the SSB phase follows phi at the true defocus and
    loses it at the wrong one (chi sign and units agree with the forward model)."""
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate, ground_truth_phase
    cfg = SyntheticConfig(scan=128, det=64, dose=1e6, seed=5, sigma_px=3.0,
                          aberrations={"C10": 200.0})
    data = generate(cfg, "cpu")
    p = O.prepare(data, voltage_kV=300, semiangle_mrad=20.0, scan_sampling=cfg.step_A,
                  det_sampling=cfg.det_sampling_mrad)
    truth = ground_truth_phase(cfg)

    def corr(c10):
        a = np.angle(O.object_wave(p.G, p.dc, p.geom, c10, 0.0, 0.0))
        a = a - a.mean()
        b = truth - truth.mean()
        return float((a * b).sum() / math.sqrt((a * a).sum() * (b * b).sum()))

    assert corr(200.0) > 0.8
    assert abs(corr(-200.0)) < 0.2


def test_detect_bf_radius_hard_disks():
    yy, xx = np.mgrid[:64, :64]
    for r0 in (10, 16, 23):
        disk = np.where((yy - 30) ** 2 + (xx - 33) ** 2 <= r0 ** 2, 100.0, 1.0).astype(np.float32)
        assert O.detect_bf_radius(disk) == ((30, 33), r0)
    small = np.where((yy - 30) ** 2 + (xx - 33) ** 2 <= 25, 100.0, 1.0).astype(np.float32)
    assert O.detect_bf_radius(small)[1] in (5, 6)
    edge = np.zeros((64, 64), np.float32)
    edge[0, 0] = 1.0
    assert O.detect_bf_radius(edge) == ((0, 0), 16)            # max_r < 2: min(shape) // 4
    with pytest.raises(ValueError, match="no positive finite"):
        O.detect_bf_radius(np.zeros((8, 8), np.float32))


def test_phase_variance_is_the_object_phase_variance():
    """The batched search's objective: with a quadratic chi it is np.var of
    the Fourier-sum object's phase (the pinned object path), and a C30 term
    enters through the full chi (a different value)."""
    data = C.poisson_data(124)
    p = O.prepare(data, bf_radius=3, **C.POISSON_KW)
    mags, angs = np.zeros(14), np.zeros(14)
    mags[0] = -120.0
    ph = np.angle(O.object_wave(p.G, p.dc, p.geom, -120.0, 0.0, 0.0))
    assert O.phase_variance(p.G, p.dc, p.geom, mags, angs) == pytest.approx(float(np.var(ph)), rel=1e-12)
    mags[4] = 4.0e5
    assert abs(O.phase_variance(p.G, p.dc, p.geom, mags, angs) - float(np.var(ph))) > 1e-3 * float(np.var(ph))


def test_phase_variance_peaks_at_the_true_defocus():
    """Contrast of a weak-phase object's SSB phase is largest at the
    generating defocus (scan 128, detector 64, C10 = 200 A)."""
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate
    cfg = SyntheticConfig(scan=128, det=64, dose=1e6, seed=5, sigma_px=3.0, aberrations={"C10": 200.0})
    data = generate(cfg, "cpu")
    p = O.prepare(data, voltage_kV=300, semiangle_mrad=20.0, scan_sampling=cfg.step_A,
                  det_sampling=cfg.det_sampling_mrad)
    grid = np.arange(0.0, 401.0, 50.0)
    var = []
    for c10 in grid:
        mags = np.zeros(14)
        mags[0] = c10
        var.append(O.phase_variance(p.G, p.dc, p.geom, mags, np.zeros(14)))
    assert grid[int(np.argmax(var))] == 200.0
