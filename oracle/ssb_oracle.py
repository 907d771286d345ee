"""CPU oracle for the single-sideband (SSB) hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` may use it, and only as the checker.

It restates, in NumPy, the algorithm of the reference CUDA SSB engine
(bobleesj/quantem.gpu, ``src/quantem/gpu/ssb/compute/cuda``).  Every function
cites the reference lines it follows; paths are relative to the reference
repository root and ``cuda/`` abbreviates ``src/quantem/gpu/ssb/compute/cuda``.

Two precisions share one code path:

* ``np.float64`` -- the FP64 oracle every GPU error is stated against;
* ``np.float32`` -- the "as shipped" FP32 restatement (FP32 storage and
  arithmetic as the reference runs it, but with accurate ``sin``/``cos``
  instead of CUDA's ``__sincosf``).  This is also the CPU baseline that
  ``bench.py`` times.

Integer work (scan padding, the uint64 mean diffraction pattern, the
bright-field mask) is exact in both modes.

Pinning (each check lives in ``tests/test_oracle.py``):

* mean DP / padding / mask -- exact integer identities and hand-built masks;
* G extraction -- textbook DFT by brute force, Hermitian symmetry, the
  reference's own ``_extract_gqk`` test vector (``tests/test_ssb_cuda_128.py:513-535``);
* geometry / probe / gamma -- closed forms: pure defocus chi = pi*lambda*|k|^2*C10,
  Gamma(-k) = -conj Gamma(k) for C10/C12, zero aberrations -> Gamma = A(k)(A(Q-k)-A(Q+k));
* object / phase / loss -- IFFT linearity (mean of per-BF objects equals the
  object of the mean), the zero-input known answer
  (``qg/swift/Tests/MetalSSBKernelsTests/MetalSSBKernelsTests.swift:37-87``),
  a brute-force double loop on tiny inputs, and FP32 vs FP64 agreement;
* chi_full -- reduces to the C10/C12 chi when only those are set, and to
  the textbook C30 term (pi/2)*lambda^3*|k|^4*C30.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# --------------------------------------------------------------------------
# Optics (qg/optics/physics.py:24-86, 177-200)
# --------------------------------------------------------------------------

_COMMON_WAVELENGTHS_A = {
    80: 0.041757160772834,
    120: 0.033492152726724,
    200: 0.025079340450548,
    300: 0.019687489006849,
    400: 0.016439434169908,
}


def wavelength_angstrom(voltage_kV: float) -> float:
    """Relativistic electron wavelength in Angstrom.

    Follows ``qg/optics/physics.py:33-86``: tabulated values at the common
    integer voltages, otherwise the relativistic de Broglie formula.
    """
    v_int = int(voltage_kV)
    if v_int == voltage_kV and v_int in _COMMON_WAVELENGTHS_A:
        return _COMMON_WAVELENGTHS_A[v_int]
    h = 6.62607015e-34
    m = 9.1093837015e-31
    e = 1.602176634e-19
    c = 299792458
    volts = voltage_kV * 1000
    gamma = 1 + (e * volts) / (2 * m * c**2)
    return h / math.sqrt(2 * m * e * volts * gamma) * 1e10


# --------------------------------------------------------------------------
# Scan normalisation (cuda/backend.py:229-269)
# --------------------------------------------------------------------------

SUPPORTED_SCAN_SIZES = (128, 256, 512, 1024)


def normalize_scan(data: np.ndarray) -> np.ndarray:
    """Pad (with the truncated mean DP) or centre-crop to a supported square scan.

    ``cuda/backend.py:229-269``: target = next of 128/256/512/1024 above the
    longest side; oversize axes are centre-cropped; undersize axes are padded
    with ``(int64 sum / N).astype(dtype)`` -- a truncating cast for integers --
    with ``pad_top = (T - H) // 2``.
    """
    H, W = int(data.shape[0]), int(data.shape[1])
    if (H, W) in tuple((s, s) for s in SUPPORTED_SCAN_SIZES):
        return data
    longest = max(H, W)
    target = 128 if longest <= 128 else 256 if longest <= 256 else 512 if longest <= 512 else 1024
    if H > target or W > target:
        r0 = max(0, (H - target) // 2)
        c0 = max(0, (W - target) // 2)
        data = data[r0:r0 + min(target, H), c0:c0 + min(target, W)]
        H, W = int(data.shape[0]), int(data.shape[1])
    if H < target or W < target:
        pad_top = (target - H) // 2
        pad_left = (target - W) // 2
        det_h, det_w = data.shape[2], data.shape[3]
        flat = data.reshape(-1, det_h * det_w)
        is_int = np.issubdtype(data.dtype, np.integer)
        acc = flat.astype(np.int64 if is_int else np.float64).sum(axis=0)
        mean = (acc.reshape(det_h, det_w).astype(np.float64) / flat.shape[0]).astype(data.dtype)
        padded = np.broadcast_to(mean[None, None], (target, target, det_h, det_w)).copy()
        padded[pad_top:pad_top + H, pad_left:pad_left + W] = data
        data = padded
    return data


# --------------------------------------------------------------------------
# Mean diffraction pattern and bright-field selection
# (cuda/backend.py:420-518)
# --------------------------------------------------------------------------

def mean_dp(data: np.ndarray, detector_gain: np.ndarray | None = None) -> tuple[np.ndarray, np.ndarray]:
    """Return (per-pixel sums, float32 mean DP).

    ``cuda/backend.py:432-443``: ``frames.sum(axis=0, dtype=uint64 for
    integer samples, float64 otherwise).astype(float32) / N`` -- the quotient
    is a float32 array divided by an integer, i.e. one correctly rounded
    float32 division -- then ``* detector_gain`` (float32) when given.
    Signed integers enter the uint64 sum sign-extended, modulo 2^64 (the
    int64 sum's two's-complement bits).
    """
    frames = data.reshape(-1, data.shape[-2], data.shape[-1])
    if np.issubdtype(frames.dtype, np.unsignedinteger):
        sums = frames.sum(axis=0, dtype=np.uint64)
    elif np.issubdtype(frames.dtype, np.integer):
        sums = frames.sum(axis=0, dtype=np.int64).view(np.uint64)
    else:
        sums = frames.sum(axis=0, dtype=np.float64)
    mean = sums.astype(np.float32) / np.float32(frames.shape[0])
    if detector_gain is not None:
        mean = mean * np.asarray(detector_gain, dtype=np.float32)
    return sums, mean.astype(np.float32)


@dataclass
class BrightField:
    rows: np.ndarray          # int32 [N_bf], raster order
    cols: np.ndarray          # int32 [N_bf]
    center: tuple[float, float]
    radius: float
    near_tie: bool            # a pixel sits within float noise of a threshold


def select_bf(mean: np.ndarray, threshold: float = 0.0,
              bf_radius: float | None = None) -> BrightField:
    """Bright-field mask from a float32 mean DP.

    ``cuda/backend.py:448-518``:

    * candidate mask ``mean > max * threshold`` in raster (``nonzero``) order;
    * ``bf_radius is None``: probe mask ``mean > mean.mean() + mean.std()``,
      centre = float32(sum of coordinates) / count, radius = sqrt(count/pi);
    * explicit radius: intensity-weighted centroid of the candidates;
    * keep ``(row - cr)^2 + (col - cc)^2 <= radius^2`` evaluated in float32.

    The reference forms ``mean.mean()`` and ``mean.std()`` with float32 GPU
    reductions whose order cannot be reproduced here; this oracle uses FP64
    and reports ``near_tie`` when any pixel lies within 1e-5 (relative) of
    that threshold, so tests can tell a true disagreement from reduction
    order.
    """
    mean = np.asarray(mean, dtype=np.float32)
    mmax = np.float32(mean.max())
    cand = mean > mmax * np.float32(threshold)
    rows, cols = np.nonzero(cand)
    rows = rows.astype(np.int32)
    cols = cols.astype(np.int32)
    if rows.size == 0:
        raise ValueError(f"No bright-field pixels found with threshold {threshold:.2f}.")
    near_tie = False
    if bf_radius is None:
        m64 = mean.astype(np.float64)
        thr = m64.mean() + m64.std()
        probe = m64 > thr
        near_tie = bool(np.any(np.abs(m64 - thr) <= 1e-5 * max(abs(thr), 1e-30)))
        total = int(probe.sum())
        if total > 0:
            rr, cc = np.nonzero(probe)
            # float32 sums of integer coordinates are exact below 2^24.
            center_row = float(np.float32(np.float32(rr.sum()) / np.float32(total)))
            center_col = float(np.float32(np.float32(cc.sum()) / np.float32(total)))
            radius = math.sqrt(total / math.pi)
        else:
            center_row = mean.shape[0] / 2.0
            center_col = mean.shape[1] / 2.0
            radius = min(mean.shape) * 0.25
    else:
        w = mean[rows, cols].astype(np.float64)
        wsum = w.sum()
        if wsum > 0:
            center_row = float(np.float32((rows.astype(np.float64) * w).sum() / wsum))
            center_col = float(np.float32((cols.astype(np.float64) * w).sum() / wsum))
        else:
            center_row = float(np.float32(rows.mean()))
            center_col = float(np.float32(cols.mean()))
        radius = float(bf_radius)
    dr = rows.astype(np.float32) - np.float32(center_row)
    dc = cols.astype(np.float32) - np.float32(center_col)
    dist_sq = dr * dr + dc * dc
    r2 = np.float32(radius * radius)
    within = dist_sq <= r2
    if np.any(np.abs(dist_sq - r2) <= 1e-5 * float(r2)):
        near_tie = True
    rows, cols = rows[within], cols[within]
    if rows.size == 0:
        raise ValueError(f"No bright-field pixels within bf_radius={radius}.")
    return BrightField(rows, cols, (center_row, center_col), radius, near_tie)


def detect_bf_radius(mean: np.ndarray, threshold_ratio: float = 0.1) -> tuple[tuple[int, int], int]:
    """Probe disk centre and radius, ``qg/detector/compute/cuda/probe.py:12-123``.

    Centre of mass of ``mean > ratio*max`` (float32 sums of integer
    coordinates, exact; rounded half-to-even), radial profile over
    ``rint(|r - centre|)`` bins below the distance to the nearest edge, a
    13-tap Gaussian (sigma 2) with edge padding, and the first bin under half
    of the mean of the first five smoothed bins.  Profile sums and the
    smoothing run in FP64 here.
    """
    dp = np.asarray(mean, dtype=np.float32)
    if dp.ndim != 2:
        raise ValueError(f"Expected 2D diffraction pattern, got {dp.ndim}D with shape {dp.shape}")
    R, C = dp.shape
    if R == 0 or C == 0:
        raise ValueError(f"Diffraction pattern has zero-size dimension: shape {dp.shape}")
    with np.errstate(all="ignore"):
        dmax = float(np.nanmax(dp))
    if not np.isfinite(dmax) or dmax <= 0:
        raise ValueError("Diffraction pattern has no positive finite values - cannot detect BF disk. "
                         "Check that your data is loaded correctly.")
    mask = dp > np.float32(threshold_ratio * dmax)
    if not mask.any():
        raise ValueError(f"No pixels above threshold ({threshold_ratio:.0%} of max intensity). "
                         "The diffraction pattern may be too noisy or empty.")
    rr, cc = np.nonzero(mask)
    total = np.float32(rr.size)
    rcf = float(np.float32(np.float32(rr.sum()) / total))
    ccf = float(np.float32(np.float32(cc.sum()) / total))
    rc = max(0, min(int(round(rcf)), R - 1))
    col_c = max(0, min(int(round(ccf)), C - 1))
    max_r = min(rc, col_c, R - rc, C - col_c)
    if max_r < 2:
        return (rc, col_c), max(1, min(R, C) // 4)
    dr = np.arange(R, dtype=np.float32)[:, None] - np.float32(rc)
    dc = np.arange(C, dtype=np.float32)[None, :] - np.float32(col_c)
    rint = np.rint(np.sqrt(dr * dr + dc * dc)).astype(np.int64).ravel()
    flat = dp.astype(np.float64).ravel()
    valid = rint < max_r
    profile = np.bincount(rint[valid], weights=flat[valid], minlength=max_r)[:max_r]
    counts = np.bincount(rint[valid], minlength=max_r)[:max_r]
    nz = counts > 0
    profile[nz] /= counts[nz]
    if profile.size > 5:
        x = np.arange(13, dtype=np.float64) - 6
        kern = np.exp(-0.5 * (x / 2.0) ** 2)
        kern /= kern.sum()
        padded = np.pad(profile, 6, mode="edge")
        smooth = np.convolve(padded, kern, mode="valid")[:profile.size]
        half = smooth[:5].mean() * 0.5
        below = np.nonzero(smooth < half)[0]
        radius = int(below[0]) if below.size else int(profile.size) // 2
    else:
        radius = min(R, C) // 4
    return (rc, col_c), max(1, radius)


# --------------------------------------------------------------------------
# Scan Fourier stack (cuda/backend.py:520-589)
# --------------------------------------------------------------------------

def extract_gqk(data: np.ndarray, rows: np.ndarray, cols: np.ndarray,
                dtype=np.float64, detector_gain: np.ndarray | None = None) -> tuple[np.ndarray, complex]:
    """Hermitian half-plane ``G[b, Qr, 0..N/2]`` and ``dc = mean_b G_b[0,0]``.

    ``cuda/backend.py:520-589``: gather BF columns, transpose to
    ``[b, R_r, R_c]``, cast to complex64 (samples rounded to float32), times
    the BF pixels' detector gain in float32 when given (``:574-576``),
    unnormalised forward ``fft2``, keep columns ``0..N/2``.  The oracle uses
    ``rfft2`` (identical columns) at the requested precision.
    """
    import scipy.fft as sfft
    H, W = data.shape[:2]
    flat = data.reshape(H * W, data.shape[2], data.shape[3])
    vbf = flat[:, rows, cols].T.reshape(len(rows), H, W).astype(np.float32)
    if detector_gain is not None:
        vbf = vbf * np.asarray(detector_gain, dtype=np.float32)[rows, cols][:, None, None]
    real = np.float32 if dtype == np.float32 else np.float64
    g = sfft.rfft2(vbf.astype(real), axes=(-2, -1), workers=_workers())
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    g = g.astype(cdt, copy=False)
    dc = complex(g[:, 0, 0].astype(np.complex128).mean())
    return g, dc


def expand_hermitian(half: np.ndarray) -> np.ndarray:
    """Full ``[b, N, N]`` plane from the stored half-plane.

    Mirror rule of ``ld_gqk_maybe_herm`` (``cuda/kernels/common.py:65-84``):
    ``G[r, c] = conj(G[(-r) % N, N - c])`` for ``c > N/2``.
    """
    nb, n, stored = half.shape
    full = np.empty((nb, n, n), dtype=half.dtype)
    full[:, :, :stored] = half
    mrows = (-np.arange(n)) % n
    mcols = np.arange(n - stored, 0, -1)
    full[:, :, stored:] = np.conj(half[:, mrows][:, :, mcols])
    return full


def _full_plane(g: np.ndarray) -> np.ndarray:
    """A full ``[b, N, N]`` stack as is; a Hermitian half-plane expanded.

    The reference engine accepts both storages (``cuda/engine.py:497-502``)."""
    return g if g.shape[2] == g.shape[1] else expand_hermitian(g)


def _workers() -> int:
    import os
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


# --------------------------------------------------------------------------
# Geometry (cuda/backend.py:44-82, 309-327; cuda/engine.py:784-848)
# --------------------------------------------------------------------------

def scan_frequencies(n: int, sampling: float, dtype=np.float64) -> np.ndarray:
    """``fftfreq(n, sampling)``; the reference stores it as float32 (backend.py:69-70)."""
    return np.fft.fftfreq(n, sampling).astype(dtype)


@dataclass
class Geometry:
    """Per-BF detector geometry and scalars of ``SSBEngine.cache_rotation``."""
    kx: np.ndarray           # row-direction k of each BF pixel [1/A]
    ky: np.ndarray           # column-direction k
    alpha2: np.ndarray
    cos2: np.ndarray
    sin2: np.ndarray
    aperture: np.ndarray
    qx: np.ndarray           # scan row frequencies
    qy: np.ndarray           # scan column frequencies
    wavelength: float
    semiangle_rad: float
    ang_y_rad: float
    ang_x_rad: float
    dtype: type


def detector_geometry(rows, cols, center, det_shape, det_sampling_mrad,
                      wavelength, semiangle_mrad, rotation_rad, scan_n,
                      scan_sampling, dtype=np.float64, recip=None) -> Geometry:
    """Rotation-dependent BF geometry, ``cuda/engine.py:784-848``.

    Detector sampling (``cuda/backend.py:316-327``): reciprocal sampling is
    ``det_sampling[mrad]*1e-3/lambda``; the engine's ``recip = 1/(sampling*gpts)``
    equals it.  ``kx = (row - center_row)*recip_y`` (row direction),
    ``ky = (col - center_col)*recip_x``; passive rotation by ``-rotation``.
    Aperture: ``clip((semiangle - alpha)/denom + 0.5, 0, 1)`` with
    ``denom = sqrt((cos(phi)*ang_y)^2 + (sin(phi)*ang_x)^2)``.

    ``recip`` overrides the reciprocal sampling, as an engine built directly
    from ``sampling``/``gpts`` does (``cuda/engine.py:792-793``).
    """
    dt = dtype
    ang_y, ang_x = det_sampling_mrad
    if recip is None:
        recip_y = ang_y * 1e-3 / wavelength
        recip_x = ang_x * 1e-3 / wavelength
        # the engine recomputes recip from sampling = 1/(recip*gpts); same value
        sampling_y = 1.0 / (recip_y * det_shape[0])
        sampling_x = 1.0 / (recip_x * det_shape[1])
        recip_y = 1.0 / (sampling_y * det_shape[0])
        recip_x = 1.0 / (sampling_x * det_shape[1])
    else:
        recip_y, recip_x = float(recip[0]), float(recip[1])
    iy = rows.astype(dt) - dt(center[0])
    ix = cols.astype(dt) - dt(center[1])
    kxa = iy * dt(recip_y)
    kya = ix * dt(recip_x)
    cos_a = math.cos(-rotation_rad)
    sin_a = math.sin(-rotation_rad)
    kx = kxa * dt(cos_a) + kya * dt(sin_a)
    ky = -kxa * dt(sin_a) + kya * dt(cos_a)
    kx = kx.astype(dt)
    ky = ky.astype(dt)
    k_mag = np.sqrt(kx * kx + ky * ky)
    phi = np.arctan2(ky, kx)
    alpha = k_mag * dt(wavelength)
    alpha2 = alpha * alpha
    cphi = np.cos(phi)
    sphi = np.sin(phi)
    denom = np.sqrt((cphi * dt(ang_y * 1e-3)) ** 2 + (sphi * dt(ang_x * 1e-3)) ** 2)
    semi = semiangle_mrad * 1e-3
    aperture = np.clip((dt(semi) - alpha) / denom + dt(0.5), 0, 1).astype(dt)
    cos2 = (cphi * cphi - sphi * sphi).astype(dt)
    sin2 = (dt(2.0) * sphi * cphi).astype(dt)
    ss = scan_sampling if isinstance(scan_sampling, (tuple, list)) else (scan_sampling, scan_sampling)
    return Geometry(kx, ky, alpha2.astype(dt), cos2, sin2, aperture,
                    scan_frequencies(scan_n, ss[0], dt),
                    scan_frequencies(scan_n, ss[1], dt),
                    float(wavelength), float(semi), float(ang_y * 1e-3),
                    float(ang_x * 1e-3), dt)


# --------------------------------------------------------------------------
# Probe, overlap and correction (cuda/engine.py:217-231, 456;
# cuda/kernels/common.py:115-236)
# --------------------------------------------------------------------------

def aperture_geometry(dx, dy, g: Geometry):
    """``compute_geometry`` of ``cuda/kernels/common.py:115-153``.

    Returns (alpha^2, cos 2phi, sin 2phi, soft aperture) of the vector (dx, dy).
    The reference's exact inner fast path (aperture == 1 inside
    ``semiangle - max_ang/2``) gives the same value as the full expression,
    so the oracle evaluates the full expression everywhere.
    """
    dt = g.dtype
    dx2 = dx * dx
    dy2 = dy * dy
    r2 = dx2 + dy2
    alpha2 = (r2 * dt(g.wavelength)) * dt(g.wavelength)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv_r2 = np.where(r2 > dt(1e-30), dt(1.0) / r2, dt(0.0))
        cos2 = (dx2 - dy2) * inv_r2
        sin2 = dt(2.0) * dx * dy * inv_r2
        r = np.sqrt(r2)
        alpha = r * dt(g.wavelength)
        denom_num2 = (dx * dt(g.ang_y_rad)) ** 2 + (dy * dt(g.ang_x_rad)) ** 2
        inv_r = np.where(r > dt(1e-15), dt(1.0) / r, dt(0.0))
        denom = np.sqrt(denom_num2) * inv_r
        edge = np.where(denom > dt(1e-15),
                        (dt(g.semiangle_rad) - alpha) / denom + dt(0.5), dt(1.0))
    aperture = np.clip(edge, dt(0.0), dt(1.0))
    return alpha2, cos2, sin2, aperture


def chi_c10c12(alpha2, cos2, sin2, c10, c12, phi12, wavelength, dtype):
    """``chi = (pi/lambda)*alpha^2*(C10 + C12*cos 2(phi - phi12))`` (``cuda/engine.py:217-231, 456``).

    Coefficients are in Angstrom, the unit the reference's arithmetic uses
    (its documentation says nm; the numbers are consumed as Angstrom).
    """
    dt = dtype
    factor = dt(math.pi / wavelength)
    c2p = dt(math.cos(2.0 * phi12))
    s2p = dt(math.sin(2.0 * phi12))
    return factor * alpha2 * (dt(c12) * (cos2 * c2p + sin2 * s2p) + dt(c10))


def probe_pk(g: Geometry, c10, c12, phi12) -> np.ndarray:
    """``P(k_b) = A(k_b) exp(-i chi(k_b))`` (``_pk_kernel``, ``cuda/engine.py:217-231``)."""
    chi = chi_c10c12(g.alpha2, g.cos2, g.sin2, c10, c12, phi12, g.wavelength, g.dtype)
    cdt = np.complex64 if g.dtype == np.float32 else np.complex128
    return (g.aperture * (np.cos(chi) - 1j * np.sin(chi))).astype(cdt)


def _normalize_gamma(gamma, dtype):
    """``|G|^2 > 1e-16 ? G*rsqrt(|G|^2) : G*1e8`` (``cuda/kernels/common.py:227-230``)."""
    dt = dtype
    mag2 = gamma.real * gamma.real + gamma.imag * gamma.imag
    with np.errstate(divide="ignore", invalid="ignore"):
        scale = np.where(mag2 > dt(1e-16), dt(1.0) / np.sqrt(mag2), dt(1e8))
    return gamma * scale.astype(dt)


def gamma_hat_block(g: Geometry, b0: int, b1: int, c10, c12, phi12) -> np.ndarray:
    """Normalised overlap ``Gamma^_b(Q)`` on the full Q grid for BF ``b0:b1``.

    ``gamma_mul_pk_onthefly`` (``cuda/kernels/common.py:173-236``):
    ``Gamma = P(Q-k) conj(P(k)) - conj(P(Q+k)) P(k)`` with on-the-fly geometry
    at ``Q -/+ k`` and the 1e-16 / 1e8 guard.  Returns shape ``[b, Nr, Nc]``.
    """
    dt = g.dtype
    cdt = np.complex64 if dt == np.float32 else np.complex128
    qx = g.qx[None, :, None]
    qy = g.qy[None, None, :]
    kx = g.kx[b0:b1, None, None]
    ky = g.ky[b0:b1, None, None]
    pk = probe_pk(g, c10, c12, phi12)[b0:b1, None, None]
    am2, cm, sm, apm = aperture_geometry(qx - kx, qy - ky, g)
    ap2, cp_, sp, app = aperture_geometry(qx + kx, qy + ky, g)
    chi_m = chi_c10c12(am2, cm, sm, c10, c12, phi12, g.wavelength, dt)
    chi_p = chi_c10c12(ap2, cp_, sp, c10, c12, phi12, g.wavelength, dt)
    pm = (apm * (np.cos(chi_m) - 1j * np.sin(chi_m))).astype(cdt)
    pp = (app * (np.cos(chi_p) - 1j * np.sin(chi_p))).astype(cdt)
    gamma = pm * np.conj(pk) - np.conj(pp) * pk
    return _normalize_gamma(gamma.astype(cdt), dt)


# --------------------------------------------------------------------------
# Full 14-coefficient chi (cuda/kernels/common.py:410-532, 625-674)
# --------------------------------------------------------------------------

_ABR_N_PLUS_ONE_INV = np.array([1/2, 1/2, 1/3, 1/3, 1/4, 1/4, 1/4,
                                1/5, 1/5, 1/5, 1/6, 1/6, 1/6, 1/6])
_ABR_M = np.array([0, 2, 1, 3, 0, 2, 4, 1, 3, 5, 0, 2, 4, 6])
_ABR_N = np.array([1, 1, 2, 2, 3, 3, 3, 4, 4, 4, 5, 5, 5, 5])
ABERRATION_NAMES = ("C10", "C12", "C21", "C23", "C30", "C32", "C34",
                    "C41", "C43", "C45", "C50", "C52", "C54", "C56")


def chi_full(dx, dy, wavelength, mags, angles, dtype=np.float64):
    """Krivanek chi, ``cuda/kernels/common.py:410-471``.

    ``chi = (2pi/lambda) * sum_i alpha^(n_i+1)/(n_i+1) * C_i * cos(m_i (phi - phi_i))``
    with ``alpha = lambda*|k|``; index layout C10, C12, C21, C23, C30, C32,
    C34, C41, C43, C45, C50, C52, C54, C56 (``cuda/backend.py:730-769``).
    Written directly from the definition (cos(m*phi - m*phi_i)) rather than
    the reference's Chebyshev recurrence.
    """
    dt = dtype
    r2 = dx * dx + dy * dy
    r = np.sqrt(r2)
    alpha = r * dt(wavelength)
    phi = np.arctan2(dy, dx)
    chi = np.zeros(np.broadcast(dx, dy).shape, dtype=dt)
    for i in range(14):
        if mags[i] == 0.0:
            continue
        n, m = int(_ABR_N[i]), int(_ABR_M[i])
        term = alpha ** (n + 1) * dt(mags[i] * _ABR_N_PLUS_ONE_INV[i])
        if m:
            term = term * np.cos(dt(m) * phi - dt(m * angles[i]))
        chi = chi + term
    return (dt(2.0 * math.pi / wavelength) * chi).astype(dt)


def gamma_hat_block_full(g: Geometry, b0: int, b1: int, mags, angles) -> np.ndarray:
    """``gamma_mul_pk_onthefly_full`` (``cuda/kernels/common.py:484-532``) with
    ``P(k) = A(k) exp(-i chi_full(k))`` (``_pk_kernel_full``, ``cuda/engine.py:250-303``)."""
    dt = g.dtype
    cdt = np.complex64 if dt == np.float32 else np.complex128
    kxb = g.kx[b0:b1]
    kyb = g.ky[b0:b1]
    chi_k = chi_full(kxb, kyb, g.wavelength, mags, angles, dt)
    pk = (g.aperture[b0:b1] * (np.cos(chi_k) - 1j * np.sin(chi_k))).astype(cdt)[:, None, None]
    qx = g.qx[None, :, None]
    qy = g.qy[None, None, :]
    kx = kxb[:, None, None]
    ky = kyb[:, None, None]
    _, _, _, apm = aperture_geometry(qx - kx, qy - ky, g)
    _, _, _, app = aperture_geometry(qx + kx, qy + ky, g)
    chi_m = chi_full(qx - kx, qy - ky, g.wavelength, mags, angles, dt)
    chi_p = chi_full(qx + kx, qy + ky, g.wavelength, mags, angles, dt)
    pm = (apm * (np.cos(chi_m) - 1j * np.sin(chi_m))).astype(cdt)
    pp = (app * (np.cos(chi_p) - 1j * np.sin(chi_p))).astype(cdt)
    gamma = pm * np.conj(pk) - np.conj(pp) * pk
    return _normalize_gamma(gamma.astype(cdt), dt)


def aberration_vectors(aberrations: dict) -> tuple[np.ndarray, np.ndarray, bool]:
    """Pack an aberration dict into the 14-slot (mags, angles) layout.

    ``CudaSSBBackend._ho_arrays_from_aberrations`` (``cuda/backend.py:730-769``):
    ``C10``/``C12``/``phi12`` (radians) fill slots 0-1; ``Cn0`` scalars and
    ``Cnm_mag``/``Cnm_angle`` (degrees) fill the rest; higher order is
    "active" when any slot >= 2 is non-zero.
    """
    mags = np.zeros(14)
    angs = np.zeros(14)
    mags[0] = float(aberrations.get("C10", 0.0))
    mags[1] = float(aberrations.get("C12", 0.0))
    angs[1] = float(aberrations.get("phi12", 0.0))
    active = False
    for idx, name in enumerate(ABERRATION_NAMES):
        if idx < 2:
            continue
        if _ABR_M[idx] == 0:
            mag, ang = float(aberrations.get(name, 0.0)), 0.0
        else:
            mag = float(aberrations.get(f"{name}_mag", 0.0))
            ang = math.radians(float(aberrations.get(f"{name}_angle", 0.0)))
        if mag != 0.0:
            active = True
        mags[idx] = mag
        angs[idx] = ang
    return mags, angs, active


# --------------------------------------------------------------------------
# Reconstructions (cuda/engine.py:1047-1191, 1244-1296, 1464-1514;
# cuda/backend.py:771-851)
# --------------------------------------------------------------------------

def _bf_chunk(n_bf: int, n: int, dtype) -> int:
    per = n * n * (16 if dtype == np.float64 else 8) * 6
    return max(1, min(n_bf, (512 << 20) // per))


def corrected_fourier_mean(G_half, dc, g: Geometry, c10, c12, phi12,
                           gamma_fn=None) -> np.ndarray:
    """``mean_b G_b(Q) conj(Gamma^_b(Q))`` with the DC bin set to ``dc``.

    The Fourier-sum object path: ``ssb*_corrected_fourier_sum_t256``
    (``cuda/kernels/fft256.py:967-1030``) writes ``count*dc`` at Q = 0 and
    ``_reconstruct_object_fourier_sum`` (``cuda/engine.py:1047-1107``)
    divides the group sum by N_bf.  Summation here is in the oracle's
    precision (pairwise NumPy sums).
    """
    from concurrent.futures import ThreadPoolExecutor
    dt = g.dtype
    cdt = np.complex64 if dt == np.float32 else np.complex128
    nb, n = G_half.shape[0], G_half.shape[1]
    acc = np.zeros((n, n), dtype=np.complex128)
    step = _bf_chunk(nb, n, dt)
    starts = list(range(0, nb, step))

    def chunk(b0):
        b1 = min(nb, b0 + step)
        gh = gamma_fn(b0, b1) if gamma_fn else gamma_hat_block(g, b0, b1, c10, c12, phi12)
        full = _full_plane(G_half[b0:b1].astype(cdt))
        return (full * np.conj(gh)).sum(axis=0)

    # chunks on a thread pool; partial sums added in chunk order
    workers = max(1, min(len(starts), _workers()))
    if workers == 1:
        for b0 in starts:
            acc += chunk(b0)
    else:
        with ThreadPoolExecutor(max_workers=workers) as ex:
            for part in ex.map(chunk, starts):
                acc += part
    acc /= nb
    acc[0, 0] = dc
    return acc.astype(cdt)


def object_wave(G_half, dc, g: Geometry, c10, c12, phi12) -> np.ndarray:
    """Complex object ``O = ifft2(mean_b corrected_b)`` (``cuda/engine.py:1047-1107``).

    Equal, by IFFT linearity, to ``mean_b ifft2(corrected_b)`` -- the
    fused-IFFT reference path (``cuda/engine.py:1128-1151``).
    """
    s = corrected_fourier_mean(G_half, dc, g, c10, c12, phi12)
    return np.fft.ifft2(s.astype(np.complex128)).astype(s.dtype)


def _per_bf_moments(G_half, dc, g: Geometry, gamma_fn):
    """(sum_b phi_b, sum_b phi_b^2) over BF chunks.  Chunks run on a thread
    pool (NumPy's elementwise work and the FFTs release the GIL) and their
    partial sums are added in chunk order, so the result does not depend on
    the thread count."""
    from concurrent.futures import ThreadPoolExecutor

    import scipy.fft as sfft
    dt = g.dtype
    cdt = np.complex64 if dt == np.float32 else np.complex128
    nb, n = G_half.shape[0], G_half.shape[1]
    step = _bf_chunk(nb, n, dt)
    starts = list(range(0, nb, step))
    workers = max(1, min(len(starts), _workers()))

    def chunk(b0):
        b1 = min(nb, b0 + step)
        gh = gamma_fn(b0, b1)
        corr = _full_plane(G_half[b0:b1].astype(cdt)) * np.conj(gh)
        corr[:, 0, 0] = dc
        ob = sfft.ifft2(corr, axes=(-2, -1), workers=1 if workers > 1 else _workers())
        ph = np.angle(ob).astype(dt)
        return ph.sum(axis=0, dtype=np.float64), (ph.astype(np.float64) ** 2).sum(axis=0)

    s1 = np.zeros((n, n), dtype=np.float64)
    s2 = np.zeros((n, n), dtype=np.float64)
    if workers == 1:
        parts = map(chunk, starts)
        for a, b in parts:
            s1 += a
            s2 += b
        return s1, s2
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for a, b in ex.map(chunk, starts):
            s1 += a
            s2 += b
    return s1, s2


def phase_and_loss(G_half, dc, g: Geometry, c10, c12, phi12) -> tuple[np.ndarray, float]:
    """Mean per-BF phase and the exact phase-variance loss.

    ``SSBEngine.reconstruct_with_loss`` (``cuda/engine.py:1464-1514``) and the
    reference restatement ``_reference_phase_loss`` (``tests/test_ssb_cuda_128.py:60-128``):
    ``O_b = ifft2(G_b conj(Gamma^_b))`` with ``O_b``'s DC input set to dc,
    ``phi_b = angle(O_b)``, ``phase = mean_b phi_b``,
    ``L = mean_R(mean_b phi_b^2 - phase^2)``.
    """
    nb = G_half.shape[0]
    s1, s2 = _per_bf_moments(G_half, dc, g,
                             lambda b0, b1: gamma_hat_block(g, b0, b1, c10, c12, phi12))
    phase = s1 / nb
    loss = float(np.mean(s2 / nb - phase * phase))
    return phase.astype(g.dtype), loss


def phase_and_loss_full(G_half, dc, g: Geometry, mags, angles) -> tuple[np.ndarray, float]:
    """Higher-order (14-coefficient) mean phase and loss,
    ``reconstruct_full(_with_loss)`` (``cuda/engine.py:1208-1458``)."""
    nb = G_half.shape[0]
    s1, s2 = _per_bf_moments(G_half, dc, g,
                             lambda b0, b1: gamma_hat_block_full(g, b0, b1, mags, angles))
    phase = s1 / nb
    loss = float(np.mean(s2 / nb - phase * phase))
    return phase.astype(g.dtype), loss


def phase_variance(G_half, dc, g: Geometry, mags, angles) -> float:
    """Phase-image variance ``var_R angle(O(R))`` of the Fourier-sum object
    ``O = ifft2(mean_b G_b conj(Gamma^_b))`` (DC set to dc) for a 14-coefficient
    aberration vector: the object path of ``cuda/engine.py:1047-1107`` with
    ``gamma_mul_pk_onthefly_full`` (``cuda/kernels/common.py:484-532``),
    reduced to the contrast measure the batched search maximises (no
    reference counterpart; the reference's default objective is
    ``phase_and_loss``'s)."""
    s = corrected_fourier_mean(G_half, dc, g, 0.0, 0.0, 0.0,
                               gamma_fn=lambda b0, b1: gamma_hat_block_full(g, b0, b1, mags, angles))
    ph = np.angle(np.fft.ifft2(s.astype(np.complex128)))
    return float(np.mean(ph * ph) - np.mean(ph) ** 2)


def result_phase(G_half, dc, g: Geometry, aberrations: dict) -> tuple[np.ndarray, str]:
    """Phase returned by ``result()`` (``cuda/backend.py:771-851``).

    Higher-order active -> mean per-BF phase (object synthesised as
    ``exp(i*phase)``); otherwise ``angle`` of the Fourier-sum object.
    Returns (phase, definition) with definition "object" or "mean-bf".
    """
    mags, angs, active = aberration_vectors(aberrations)
    if active:
        phase, _ = phase_and_loss_full(G_half, dc, g, mags, angs)
        return phase, "mean-bf"
    obj = object_wave(G_half, dc, g, mags[0], mags[1], angs[1])
    return np.angle(obj).astype(g.dtype), "object"


# --------------------------------------------------------------------------
# End-to-end plain SSB (the R1 boundary) and the BF column export
# --------------------------------------------------------------------------

@dataclass
class Prepared:
    data: np.ndarray
    sums: np.ndarray
    mean: np.ndarray
    bf: BrightField
    G: np.ndarray
    dc: complex
    geom: Geometry


def prepare(data, *, voltage_kV, semiangle_mrad, scan_sampling, det_sampling,
            rotation_rad=0.0, bf_intensity_threshold=0.0, bf_radius=None,
            dtype=np.float64, detector_gain=None) -> Prepared:
    """``CudaSSBBackend.__init__`` data path (``cuda/backend.py:168-365``):
    scan normalisation, mean DP, BF mask, G extraction, geometry."""
    data = normalize_scan(np.asarray(data))
    sums, mean = mean_dp(data, detector_gain)
    bf = select_bf(mean, bf_intensity_threshold, bf_radius)
    G, dc = extract_gqk(data, bf.rows, bf.cols, dtype, detector_gain)
    ds = det_sampling if isinstance(det_sampling, (tuple, list)) else (det_sampling, det_sampling)
    lam = wavelength_angstrom(voltage_kV)
    geom = detector_geometry(bf.rows, bf.cols, bf.center, data.shape[2:], ds, lam,
                             semiangle_mrad, rotation_rad, data.shape[0], scan_sampling, dtype)
    return Prepared(data, sums, mean, bf, G, dc, geom)


def gather_bf_columns(data: np.ndarray, rows: np.ndarray, cols: np.ndarray) -> np.ndarray:
    """Exact detector-major BF columns ``[b, scan]`` in the source dtype.

    ``SSBEngine.write_exact_bf_source`` / ``gather_transpose``
    (``cuda/engine.py:577-701``; checked at ``tests/test_ssb_cuda_128.py:291-303``):
    ``data.reshape(-1, kr, kc)[:, rows, cols].T``.
    """
    flat = data.reshape(-1, data.shape[-2], data.shape[-1])
    return np.ascontiguousarray(flat[:, rows, cols].T)


def bin_sum(data: np.ndarray, factor: int, pixel_mask: np.ndarray | None = None, dtype=None) -> np.ndarray:
    """Integer detector binning, ``_bin_sum`` (``qg/io/backends/cpu/reference.py:27-42``):
    uint64 sums of ``factor x factor`` blocks, incomplete edges trimmed.
    ``pixel_mask`` (nonzero = dead) zeroes those pixels first
    (``reference.py:96-111``); ``dtype`` casts the sums as ``astype`` does
    (the CPU loader casts back to the source dtype, keeping the low bits;
    the CUDA loader widens binned loads to uint32, ``qg/io/load.py:3009-3026``).
    Default: the uint64 sums."""
    if pixel_mask is not None:
        data = np.where(np.asarray(pixel_mask) != 0, 0, data).astype(data.dtype)
    kr, kc = data.shape[-2] // factor, data.shape[-1] // factor
    trimmed = data[..., :kr * factor, :kc * factor]
    shaped = trimmed.reshape(*data.shape[:-2], kr, factor, kc, factor)
    summed = shaped.astype(np.uint64).sum(axis=(-3, -1))
    return summed if dtype is None else summed.astype(dtype)


__all__ = [name for name in dir() if not name.startswith("_")]
