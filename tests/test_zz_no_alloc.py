"""The allocation-free scan normalisation, ``ssb_normalize_scan_u16_noalloc``,
against ``ssb_normalize_scan_u16`` through the C ABI: bit-identical outputs
on two calls in a row (the second with a smaller input, so stale scratch
contents from the first would show), and a clear error, before any launch,
for a workspace beyond the scratch limit."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _normalize(torch, name, src):
    from quantem.gpu_amd.ssb.compute.hip import _abi
    H, W, kr, kc = (int(v) for v in src.shape)
    T = int(_abi.lib().ssb_scan_target(H, W))
    out = torch.full((T, T, kr, kc), 7, dtype=torch.uint16, device=src.device)
    stream = _abi.stream_handle(torch.cuda.current_stream(src.device))
    _abi.call(name, _abi.ptr(src), H, W, kr, kc, _abi.ptr(out), T, stream)
    return out


def test_noalloc_normalisation_matches_the_allocating_one(cuda):
    torch = cuda
    rng = np.random.default_rng(17)
    first = torch.from_numpy(rng.integers(0, 65536, size=(200, 230, 32, 32), dtype=np.uint16)).cuda()
    second = torch.from_numpy(rng.integers(0, 4096, size=(60, 90, 16, 16), dtype=np.uint16)).cuda()
    for src in (first, second):
        ref = _normalize(torch, "ssb_normalize_scan_u16", src)
        got = _normalize(torch, "ssb_normalize_scan_u16_noalloc", src)
        torch.cuda.synchronize()
        assert torch.equal(got, ref)
    # a crop (no pad frame) and a native size pass through too
    crop = torch.from_numpy(rng.integers(0, 65536, size=(300, 140, 8, 8), dtype=np.uint16)).cuda()
    assert torch.equal(_normalize(torch, "ssb_normalize_scan_u16_noalloc", crop),
                       _normalize(torch, "ssb_normalize_scan_u16", crop))


def test_noalloc_rejects_workspace_beyond_the_limit(cuda):
    from quantem.gpu_amd.ssb.compute.hip import _abi
    torch = cuda
    src = torch.zeros((1, 1, 1, 1), dtype=torch.uint16, device="cuda")
    dst = torch.zeros((1, 1, 1, 1), dtype=torch.uint16, device="cuda")
    lib = _abi.lib()
    # 1000 x 1000 scan (16 row groups) with a 256 x 512 detector: 8 MiB of partials
    rc = lib.ssb_normalize_scan_u16_noalloc(_abi.ptr(src), 1000, 1000, 256, 512, _abi.ptr(dst), 1024,
                                            ctypes.c_void_p(0))
    assert rc == _abi.QSSB_ERR_ARG
    assert "bytes of scratch" in _abi.last_error()
