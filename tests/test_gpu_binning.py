"""Integer detector binning on the GPU (``ssb_bin_u16``) against the
oracle's ``bin_sum`` (the reference's ``_bin_sum``, dead pixels zeroed
first; uint32 = the widened binned load, uint16 = the cast back to the
source dtype): bit-exact, including trimmed edges and 16-bit wrap."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import ssb_oracle as O

pytestmark = pytest.mark.gpu


def _bin(t, factor, mask=None, dtype=None, out=None):
    from quantem.gpu_amd.ssb.compute.hip.binning import bin_frames
    return bin_frames(t, factor, pixel_mask=mask, dtype=dtype, out=out)


@pytest.mark.parametrize("factor", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", [(5, 16, 16), (2, 3, 10, 11), (1, 128, 128)])
@pytest.mark.parametrize("masked", [False, True])
def test_binning_is_bit_exact(cuda, factor, shape, masked):
    torch = cuda
    rng = np.random.default_rng(factor * 100 + len(shape))
    data = rng.integers(0, 65536, size=shape, dtype=np.uint16)
    data.reshape(-1, *shape[-2:])[0, :factor, :factor] = 65535          # forces a 16-bit wrap
    mask = (rng.random(shape[-2:]) < 0.1) if masked else None
    if mask is not None:
        mask[:factor, :factor] = False
    t = torch.from_numpy(data).cuda()
    print(f"binning factor {factor} shape {shape} masked {masked}", flush=True)
    got32 = _bin(t, factor, mask).cpu().numpy()
    got16 = _bin(t, factor, mask, torch.uint16).cpu().numpy()
    np.testing.assert_array_equal(got32, O.bin_sum(data, factor, mask, np.uint32))
    np.testing.assert_array_equal(got16, O.bin_sum(data, factor, mask, np.uint16))
    assert got32.dtype == np.uint32 and got16.dtype == np.uint16


def test_binning_benchmark_frames_by_two(cuda):
    """The benchmark's 65,536 frames of 128x128 binned by 2: the first 512
    frames bit-exact against the oracle, and every frame's binned total
    equals its raw total (no dead pixels, no trimmed edges)."""
    torch = cuda
    from quantem.gpu_amd.ssb.synthetic import SyntheticConfig, generate
    cfg = SyntheticConfig(scan=256, det=128, aberrations={"C10": 200.0}, dose=5e4, seed=2026)
    data = generate(cfg, device="cuda")
    print("binning the 256x256x128x128 benchmark frames by 2", flush=True)
    binned = _bin(data, 2)
    assert tuple(binned.shape) == (256, 256, 64, 64)
    head = data.reshape(-1, 128, 128)[:512].cpu().numpy()
    np.testing.assert_array_equal(binned.reshape(-1, 64, 64)[:512].cpu().numpy(), O.bin_sum(head, 2, None, np.uint32))
    raw_tot = data.reshape(65536, -1).to(torch.int64).sum(dim=1)
    bin_tot = binned.reshape(65536, -1).to(torch.int64).sum(dim=1)
    assert torch.equal(raw_tot, bin_tot)


def test_binning_out_buffer_and_errors(cuda):
    torch = cuda
    t = torch.from_numpy(np.full((2, 8, 8), 7, dtype=np.uint16)).cuda()
    out = torch.empty((2, 4, 4), dtype=torch.uint32, device="cuda")
    assert _bin(t, 2, out=out) is out
    np.testing.assert_array_equal(out.cpu().numpy(), np.full((2, 4, 4), 28, dtype=np.uint32))
    with pytest.raises(ValueError):
        _bin(t, 0)
    with pytest.raises(ValueError):
        _bin(t, 9)
    with pytest.raises(ValueError):
        _bin(t, 2, mask=np.zeros((4, 4)))
    with pytest.raises(ValueError):
        _bin(t, 2, out=torch.empty((2, 4, 4), dtype=torch.uint16, device="cuda"))
    with pytest.raises(NotImplementedError):
        _bin(t.to(torch.int32), 2)
