"""Seeded synthetic 4D-STEM datasets for SSB (weak-phase forward model).

The reference ships no synthetic generator for its SSB path; this module is
the data contract of the benchmark configurations:

* object phase: seeded white noise on a periodic grid, Gaussian-smoothed
  (sigma in object pixels), scaled to a target RMS in radians;
* probe: inverse FFT of ``A(k) exp(-i chi(k))`` on the detector grid, with
  ``chi`` in the numeric convention of the SSB engine (C10/C12/C30 in
  Angstrom, ``chi = pi*lambda*|k|^2*(C10 + C12 cos 2(phi-phi12)) +
  (pi/2)*lambda^3*|k|^4*C30``);
* intensities ``|FFT(probe(r - R) exp(i phi(r)))|^2`` (DC at the detector
  centre), normalised to ``dose`` electrons per probe, Poisson-sampled from
  ``seed + 1`` and clipped to uint16.

Generation runs through PyTorch FFTs in batches of scan positions; it works
on CPU for small cases and on the GPU for the full-size benchmark inputs.
It is input preparation only and is never inside a timed region.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from ..optics import electron_wavelength_angstrom


@dataclass(frozen=True)
class SyntheticConfig:
    scan: int = 64
    det: int = 64
    voltage_kV: float = 300.0
    semiangle_mrad: float = 20.0
    det_sampling_mrad: float = 1.25
    scan_step_A: float | None = None       # default: 1/(64*dk) = Nyquist for 2*alpha/lambda
    aberrations: dict = field(default_factory=dict)   # C10, C12, phi12 (rad), C30 (Angstrom)
    rotation_deg: float = 0.0
    dose: float = 1e4
    phase_rms: float = 0.1
    sigma_px: float = 2.0
    seed: int = 0

    @property
    def wavelength(self) -> float:
        return electron_wavelength_angstrom(self.voltage_kV * 1e3)

    @property
    def dk(self) -> float:
        """Detector reciprocal sampling [1/Angstrom]."""
        return self.det_sampling_mrad * 1e-3 / self.wavelength

    @property
    def pixel_A(self) -> float:
        """Object (probe window) pixel size [Angstrom]."""
        return 1.0 / (self.det * self.dk)

    @property
    def step_A(self) -> float:
        return self.scan_step_A if self.scan_step_A is not None else 1.0 / (64 * self.dk)

    @property
    def step_px(self) -> float:
        return self.step_A / self.pixel_A

    @property
    def object_px(self) -> int:
        return int(round(self.scan * self.step_px))


def object_phase(cfg: SyntheticConfig) -> np.ndarray:
    """Periodic smoothed-noise phase on the object grid (float64, radians)."""
    n = cfg.object_px
    rng = np.random.default_rng(cfg.seed)
    noise = rng.standard_normal((n, n))
    f = np.fft.fftfreq(n)
    kr, kc = np.meshgrid(f, f, indexing="ij")
    kernel = np.exp(-2.0 * math.pi**2 * cfg.sigma_px**2 * (kr**2 + kc**2))
    smooth = np.fft.ifft2(np.fft.fft2(noise) * kernel).real
    smooth -= smooth.mean()
    rms = float(np.sqrt(np.mean(smooth**2)))
    return smooth * (cfg.phase_rms / rms)


def probe_chi(kr, kc, cfg: SyntheticConfig):
    """Aberration phase of the probe on reciprocal coordinates (row, col)."""
    ab = cfg.aberrations
    lam = cfg.wavelength
    c10 = float(ab.get("C10", 0.0))
    c12 = float(ab.get("C12", 0.0))
    phi12 = float(ab.get("phi12", 0.0))
    c30 = float(ab.get("C30", 0.0))
    k2 = kr * kr + kc * kc
    phi = np.arctan2(kc, kr)
    chi = math.pi * lam * k2 * (c10 + c12 * np.cos(2.0 * (phi - phi12)))
    chi = chi + 0.5 * math.pi * lam**3 * k2 * k2 * c30
    return chi


def probe_fourier(cfg: SyntheticConfig) -> np.ndarray:
    """``A(k) exp(-i chi(k))`` on the unshifted detector grid (complex128)."""
    n = cfg.det
    f = np.fft.fftfreq(n, d=cfg.pixel_A)
    kr, kc = np.meshgrid(f, f, indexing="ij")
    alpha = np.sqrt(kr**2 + kc**2) * cfg.wavelength
    semi = cfg.semiangle_mrad * 1e-3
    ang = cfg.det_sampling_mrad * 1e-3
    aperture = np.clip((semi - alpha) / ang + 0.5, 0.0, 1.0)
    return aperture * np.exp(-1j * probe_chi(kr, kc, cfg))


def scan_positions(cfg: SyntheticConfig) -> np.ndarray:
    """Probe positions [scan, scan, 2] in object pixels (row, col), rotated."""
    idx = np.arange(cfg.scan) * cfg.step_px
    rr, cc = np.meshgrid(idx, idx, indexing="ij")
    # sign chosen so that a dataset generated at +rotation is reconstructed
    # at +rotation by the engine's passive detector rotation
    t = -math.radians(cfg.rotation_deg)
    ct, st = math.cos(t), math.sin(t)
    pr = rr * ct - cc * st
    pc = rr * st + cc * ct
    return np.stack([pr, pc], axis=-1)


def generate(cfg: SyntheticConfig, device: str = "cpu", batch: int = 4096,
             return_phase: bool = False):
    """Generate ``I[scan, scan, det, det]`` as uint16 (and optionally the phase).

    Returns a NumPy array when ``device == "cpu"``; otherwise a torch tensor
    resident on ``device``.
    """
    import torch

    n = cfg.det
    phase = object_phase(cfg)
    nobj = phase.shape[0]
    dev = torch.device(device)
    obj = torch.exp(1j * torch.as_tensor(phase, dtype=torch.float64, device=dev)).to(torch.complex64)
    pf = torch.as_tensor(probe_fourier(cfg), dtype=torch.complex128, device=dev)
    norm = float((pf.abs() ** 2).sum().item())
    pos = torch.as_tensor(scan_positions(cfg).reshape(-1, 2), dtype=torch.float64, device=dev)
    base = torch.floor(pos)
    frac = pos - base
    f = torch.fft.fftfreq(n, device=dev, dtype=torch.float64)
    win = torch.arange(n, device=dev) - n // 2
    total = pos.shape[0]
    out = torch.empty((total, n, n), dtype=torch.float32, device=dev)
    for s in range(0, total, batch):
        e = min(total, s + batch)
        fr = frac[s:e]
        # sub-pixel shift of the probe centred in the window
        ramp = torch.exp(-2j * math.pi * (fr[:, 0, None, None] * f[None, :, None]
                                          + fr[:, 1, None, None] * f[None, None, :]))
        probe = torch.fft.ifft2(pf[None] * ramp)
        probe = torch.fft.fftshift(probe, dim=(-2, -1)).to(torch.complex64)
        b = base[s:e].to(torch.int64)
        rows = (b[:, 0, None] + win[None, :]) % nobj
        cols = (b[:, 1, None] + win[None, :]) % nobj
        patch = obj[rows[:, :, None], cols[:, None, :]]
        exitw = probe * patch
        amp = torch.fft.fftshift(torch.fft.fft2(exitw), dim=(-2, -1))
        out[s:e] = (amp.real**2 + amp.imag**2) * (cfg.dose / norm)
    gen = torch.Generator(device=dev)
    gen.manual_seed(cfg.seed + 1)
    if out.numel() < 2 ** 31:
        counts = torch.poisson(out, generator=gen).clamp_(0, 65535)
        data = counts.to(torch.int32).to(torch.uint16).reshape(cfg.scan, cfg.scan, n, n)
    else:
        # torch's Poisson sampler launches one grid over the whole tensor;
        # above 2^31 elements (scan 512 x detector 128) sample in slabs of
        # frames, in order, from the same generator
        data = torch.empty((total, n, n), dtype=torch.uint16, device=dev)
        slab = (2 ** 30) // (n * n)
        for s in range(0, total, slab):
            e = min(total, s + slab)
            data[s:e] = torch.poisson(out[s:e], generator=gen).clamp_(0, 65535).to(torch.int32).to(torch.uint16)
        del out
        data = data.reshape(cfg.scan, cfg.scan, n, n)
    if device == "cpu":
        data = data.numpy()
    if return_phase:
        return data, phase
    return data


def ground_truth_phase(cfg: SyntheticConfig) -> np.ndarray:
    """Object phase sampled at the scan positions (nearest object pixel)."""
    phase = object_phase(cfg)
    pos = np.rint(scan_positions(cfg)).astype(np.int64) % phase.shape[0]
    return phase[pos[..., 0], pos[..., 1]]
