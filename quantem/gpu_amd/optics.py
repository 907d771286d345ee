"""Electron optics constants used by the SSB path.

Mirrors ``qg/optics/physics.py:24-86, 177-200`` of the reference: tabulated
wavelengths at the common integer voltages, otherwise the relativistic de
Broglie formula ``lambda = h / sqrt(2 m e V (1 + eV / (2 m c^2)))``.
"""
from __future__ import annotations

import math

PLANCK_H = 6.62607015e-34
ELECTRON_MASS = 9.1093837015e-31
ELECTRON_CHARGE = 1.602176634e-19
SPEED_OF_LIGHT = 299792458

COMMON_VOLTAGES = {
    80: 0.041757160772834,
    120: 0.033492152726724,
    200: 0.025079340450548,
    300: 0.019687489006849,
    400: 0.016439434169908,
}


def wavelength_A_from_kV(voltage_kV: float) -> float:
    """Relativistic electron wavelength [Angstrom] at ``voltage_kV``."""
    v_int = int(voltage_kV)
    if v_int == voltage_kV and v_int in COMMON_VOLTAGES:
        return COMMON_VOLTAGES[v_int]
    volts = voltage_kV * 1000
    gamma = 1 + (ELECTRON_CHARGE * volts) / (2 * ELECTRON_MASS * SPEED_OF_LIGHT**2)
    lam = PLANCK_H / math.sqrt(2 * ELECTRON_MASS * ELECTRON_CHARGE * volts * gamma)
    return lam * 1e10


def electron_wavelength_angstrom(energy_ev: float) -> float:
    """Wavelength [Angstrom] for an electron energy in eV."""
    return wavelength_A_from_kV(energy_ev / 1000.0)
