"""GPU-side helpers shared by the parity tests: engines built on the HIP
plan from the reference's fixtures or from raw data, and error measures."""
from __future__ import annotations

import numpy as np

from oracle import ssb_oracle as O
from tests import _cases as C


def _lib_is_in_tree():
    from quantem.gpu_amd.ssb.compute.hip import _abi
    assert _abi.lib()._name == str(_abi.LIB_PATH)


def make_engine(torch, n, det, max_bf, wavelength, semiangle_mrad, ang_mrad, step):
    from quantem.gpu_amd.ssb.compute.hip.engine import HipSSBEngine, PlanSpec, SSBPlan
    spec = PlanSpec(n, det[0], det[1], max_bf, wavelength, semiangle_mrad * 1e-3,
                    ang_mrad[0] * 1e-3, ang_mrad[1] * 1e-3, step, step)
    return HipSSBEngine(SSBPlan(spec, device=torch.device("cuda", 0)))


def engine_for_case(torch, case: C.EngineCase, g=None):
    """The reference ``_make_engine`` setup on the HIP plan (G imported)."""
    g = case.G if g is None else g
    eng = make_engine(torch, case.n, case.det, len(case.rows), case.wavelength,
                      case.semiangle_mrad, case.ang_mrad, case.q_step)
    if case.recip is not None:
        eng.set_detector_sampling(*case.recip)
    eng.set_bf(case.rows, case.cols, case.center, case.dc)
    eng.import_g(torch.from_numpy(np.ascontiguousarray(g)))
    eng.cache_rotation(0.0)
    return eng


def object_errors(obj, ref):
    err = np.abs(obj.astype(np.complex128) - ref)
    rel = err / np.maximum(np.abs(ref), 1e-6)
    return float(np.percentile(err, 99.9)), float(np.percentile(rel, 99.9))


def scaled_error(obj, ref):
    """p99.9 of |obj - ref| over the RMS of |ref|.

    The reference's absolute bars (5e-9, 1e-7) compare two storages through
    identical float32 arithmetic; across different summation orders the
    float32 bar scales with the object, so it is stated relative to its RMS:
    2e-6 is 34 float32 ulps, and at the _make_engine scale (RMS ~4e-3) it is
    the reference's 5e-9."""
    ref = np.asarray(ref, np.complex128)
    err = np.abs(np.asarray(obj, np.complex128) - ref)
    return float(np.percentile(err, 99.9) / np.sqrt(np.mean(np.abs(ref) ** 2)))


OBJ_SCALED = 2e-6


def engine_for_data(torch, data_np, *, semiangle_mrad, det_sampling, step, voltage_kV=300,
                    max_bf=4096):
    n, det = data_np.shape[0], data_np.shape[2:]
    lam = O.wavelength_angstrom(voltage_kV)
    return make_engine(torch, n, det, max_bf, lam, semiangle_mrad, (det_sampling, det_sampling), step)


def synthetic(scan, det, aberrations, rotation_deg=0.0, dose=5e4, seed=2026):
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate
    cfg = SyntheticConfig(scan=scan, det=det, aberrations=aberrations, rotation_deg=rotation_deg,
                          dose=dose, seed=seed)
    return cfg, generate(cfg, device="cuda").cpu().numpy()
