"""Write ``reference_vectors.json``: the reference's own test vectors and
known answers for the SSB path, with expected values from closed forms.

Nothing here runs or imports the reference or the oracle.  Inputs are the
ones the reference's tests construct; expected outputs come from the
mathematics (sums of geometric/arithmetic series, the DFT of a ramp, the
zero-input known answer) or are constants the reference's sources hold as
data.  Run from the repository root:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "reference_vectors.json"


def ramp_dft(n: int, a: float, b: float, off: float) -> np.ndarray:
    """Closed-form 2D DFT of ``f[r, c] = a*r + b*c + off`` on an n x n grid.

    sum_r z^r = 0 and sum_r r z^r = n/(z - 1) for z = exp(-2 pi i u/n), u != 0,
    so only the first row and column are non-zero.
    """
    out = np.zeros((n, n), dtype=np.complex128)
    tri = n * (n - 1) / 2.0
    out[0, 0] = n * n * off + a * n * tri + b * n * tri
    for u in range(1, n):
        z = np.exp(-2j * math.pi * u / n)
        out[u, 0] = a * n * n / (z - 1.0)
        out[0, u] = b * n * n / (z - 1.0)
    return out


def extract_gqk_arange() -> dict:
    """``tests/test_ssb_cuda_128.py:513-535``: ``arange(8*8*6*6)`` as uint16
    reshaped (8, 8, 6, 6), BF rows (2, 2, 3, 3), cols (2, 3, 2, 3).

    The virtual BF image of pixel (i, j) is ``(r*8 + c)*36 + i*6 + j``,
    a ramp with a = 288, b = 36, off = 6i + j.
    """
    rows = [2, 2, 3, 3]
    cols = [2, 3, 2, 3]
    half = []
    for i, j in zip(rows, cols):
        g = ramp_dft(8, 288.0, 36.0, 6.0 * i + j)
        half.append(g[:, :5])
    half = np.stack(half)
    dc = complex(half[:, 0, 0].mean())
    return {
        "source": "tests/test_ssb_cuda_128.py:513-535",
        "scan": [8, 8], "det": [6, 6], "rows": rows, "cols": cols,
        "g_half_real": half.real.tolist(), "g_half_imag": half.imag.tolist(),
        "g_half_shape": list(half.shape), "dc": [dc.real, dc.imag],
    }


def scan_ramp_128() -> dict:
    """The same ramp structure at a scan size the GPU path supports:
    ``I[r, c, i, j] = r + 2c + 8i + j`` on scan 128, det 8 (max 529 < 2^16)."""
    rows = [3, 3, 4, 4, 5]
    cols = [3, 4, 3, 4, 2]
    n = 128
    dcs = []
    first_row, first_col = None, None
    for i, j in zip(rows, cols):
        g = ramp_dft(n, 1.0, 2.0, 8.0 * i + j)
        dcs.append([g[0, 0].real, g[0, 0].imag])
        if first_row is None:
            first_row = g[0, : n // 2 + 1]
            first_col = g[:, 0]
    return {
        "recipe": "I[r,c,i,j] = r + 2*c + 8*i + j (uint16), scan 128, det 8x8",
        "scan": [n, n], "det": [8, 8], "rows": rows, "cols": cols,
        "a": 1.0, "b": 2.0, "off_row": 8.0, "off_col": 1.0,
        "g00": dcs,
        "g_row0_minus_dc_real": (first_row.real[1:]).tolist(),
        "g_row0_minus_dc_imag": (first_row.imag[1:]).tolist(),
        "g_col0_minus_dc_real": (first_col.real[1:]).tolist(),
        "g_col0_minus_dc_imag": (first_col.imag[1:]).tolist(),
    }


def zero_input_known_answer() -> dict:
    """``qg/swift/Tests/MetalSSBKernelsTests/MetalSSBKernelsTests.swift:37-87``:
    one BF pixel, zero data, dc = (1, 0), zero aberrations."""
    return {
        "source": "qg/swift/Tests/MetalSSBKernelsTests/MetalSSBKernelsTests.swift:37-87",
        "num_bf": 1, "dc": [1.0, 0.0],
        "fourier_sum": "delta at Q = 0 with value 1 (tolerance 1e-6)",
        "object_value": "1/N^2 everywhere, imaginary 0 (tolerance 2e-7)",
        "loss": 0.0, "loss_atol": 1e-8,
    }


def wavelengths() -> dict:
    """Tabulated electron wavelengths, ``qg/optics/physics.py:24-30``, and the
    CODATA constants of the relativistic formula (``:33-86``)."""
    return {
        "source": "qg/optics/physics.py:24-30",
        "table_A": {"80": 0.041757160772834, "120": 0.033492152726724,
                    "200": 0.025079340450548, "300": 0.019687489006849,
                    "400": 0.016439434169908},
        "codata": {"h": 6.62607015e-34, "m_e": 9.1093837015e-31,
                   "e": 1.602176634e-19, "c": 299792458.0},
    }


def fixtures() -> dict:
    """Construction recipes of the reference's seeded fixtures (inputs only)."""
    return {
        "make_engine": {
            "source": "tests/test_ssb_cuda_128.py:212-249",
            "seed": 1234, "row_pattern": [13, 14, 15, 16, 17, 16, 15],
            "col_pattern": [14, 15, 16, 17, 16, 15, 14], "bf_center": [15.5, 15.5],
            "gpts": [32, 32], "sampling": [1.0, 1.0], "q_step": 0.5,
            "wavelength": 0.0197, "semiangle_mrad": 21.4, "angular_sampling_mrad": [1.0, 1.0],
            "aberrations": {"C10": -120.0, "C12": 55.0, "phi12_deg": 17.0},
            "sizes": [[128, 7], [256, 5], [512, 5]],
        },
        "hermitian_stack": {"source": "tests/test_ssb_cuda_128.py:481-510", "seed": 4321},
        "poisson": {
            "source": "tests/test_ssb_cuda_128.py:587-720",
            "seeds": [123, 124, 125], "lam": 4.0, "shape": [128, 128, 16, 16],
            "disk_center": [8, 8], "disk_radius": 4, "disk_add": 80,
            "voltage_kV": 300, "semiangle_mrad": 21.4, "scan_sampling": 0.5,
            "det_sampling": 1.0, "bf_radius": 3,
            "aberrations": {"C10": -120.0, "C12": 55.0, "phi12_deg": 17.0},
            "object_abs_p999": 1e-7, "object_rel_p999": 1e-4, "phase_abs_p999": 3e-4,
        },
        "roi96": {"source": "tests/test_ssb_cuda_128.py:951-973", "seed": 5,
                  "shape": [96, 96, 16, 16], "disk_add": 100, "padded_scan": 128},
        "tolerances": {"source": "tests/ssb_precision.py:7-10",
                       "phase_rtol": 2e-4, "phase_atol": 2e-4,
                       "loss_rtol": 1e-5, "loss_atol": 1e-6,
                       "fourier_sum_abs_p999": 5e-9, "fourier_sum_rel_p999": 1e-4,
                       "phase_p999": 3e-4},
    }


FIXTURE_H5 = Path("/root/reference/src/quantem/gpu/swift/Tests/Native4DSTEMIOTests/Fixtures/fixture_data_000001.h5")
CHUNK_OUT = Path(__file__).resolve().parent / "bslz4_fixture_chunk.bin"


def bslz4_fixture_chunk() -> dict:
    """The one bitshuffle-LZ4 chunk of the reference's native-IO fixture
    (``src/quantem/gpu/swift/Tests/Native4DSTEMIOTests/Fixtures/fixture_data_000001.h5``):
    one 64 x 64 uint16 frame whose decoded values are ``pixel % 257``
    (``Native4DSTEMIOTests.swift:1224-1226``, frame read back at ``:1286``;
    the bad-pixel zeroing there happens after the decode).  The chunk is
    located by a plain byte search for its header -- total bytes 8192 and
    block bytes 8192 as big-endian u64/u32 -- and its length is the header
    plus one 4-byte block size plus that block; no HDF5 library is used."""
    import struct
    raw = FIXTURE_H5.read_bytes()
    header = struct.pack(">QI", 64 * 64 * 2, 8192)
    at = raw.find(header)
    if at < 0 or raw.find(header, at + 1) >= 0:
        raise RuntimeError("expected exactly one chunk header in the fixture")
    (block,) = struct.unpack(">I", raw[at + 12:at + 16])
    chunk = raw[at:at + 16 + block]
    CHUNK_OUT.write_bytes(chunk)
    return {"source": "src/quantem/gpu/swift/Tests/Native4DSTEMIOTests/Fixtures/fixture_data_000001.h5",
            "file": CHUNK_OUT.name, "offset": at, "bytes": len(chunk), "lz4_block_bytes": block,
            "shape": [64, 64], "dtype": "uint16", "expected": "pixel % 257 (raster order)",
            "expected_from": "Native4DSTEMIOTests.swift:1224-1226, 1286"}


def main() -> None:
    doc = {
        "extract_gqk_arange": extract_gqk_arange(),
        "scan_ramp_128": scan_ramp_128(),
        "zero_input": zero_input_known_answer(),
        "wavelength": wavelengths(),
        "fixtures": fixtures(),
        "bslz4_fixture_chunk": bslz4_fixture_chunk(),
    }
    OUT.write_text(json.dumps(doc, indent=1) + "\n")
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
