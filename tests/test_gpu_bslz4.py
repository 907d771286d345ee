"""GPU bitshuffle-LZ4 decode against the CPU oracle's encoder: bit-exact
frames over the chunk-layout edge cases, malformed chunks rejected, and a
decoded dataset reconstructing exactly like the raw one."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from oracle import bslz4_oracle as B
from tests import _cases as C
from tests._gpu import engine_for_data

pytestmark = pytest.mark.gpu


def _frames(kind, n, shape, seed):
    rng = np.random.default_rng(seed)
    size = (n,) + shape
    if kind == "disk":                  # a bright disk of radius 16 on a dark detector
        yy, xx = np.mgrid[:shape[0], :shape[1]]
        inside = (yy - shape[0] / 2) ** 2 + (xx - shape[1] / 2) ** 2 < 16 ** 2
        return (rng.poisson(60, size) * inside).astype(np.uint16)
    if kind.startswith("runs"):         # every bit-plane in random runs of whole 4- or 8-byte words
        g = int(kind[4:])
        m = int(np.prod(shape))
        v = np.zeros((n, m), np.uint32)
        for b in range(16):
            v |= np.repeat(rng.integers(0, 2, (n, m // g)), g, axis=1).astype(np.uint32) << b
        return v.reshape(size).astype(np.uint16)
    if kind == "periodic":              # bit-planes that repeat every 2 to 30 bytes
        p = np.arange(int(np.prod(shape)))
        v = ((p % 24) * 2749 + (p % 16) * 17 + (p % 40) * 3) % 65536
        return np.broadcast_to(v.reshape(shape), size).astype(np.uint16)
    return {"poisson": lambda: rng.poisson(0.8, size), "noisy": lambda: rng.poisson(20, size),
            "random": lambda: rng.integers(0, 65536, size),
            "zeros": lambda: np.zeros(size), "max": lambda: np.full(size, 65535)}[kind]().astype(np.uint16)


@pytest.mark.parametrize("kind,n,shape,per_chunk", [
    ("poisson", 6, (64, 64), 1),        # one full block per frame
    ("poisson", 3, (96, 96), 1),        # full blocks + a last block
    ("random", 5, (7, 7), 1),           # 48-element last block + 1 raw element; unaligned chunks
    ("poisson", 4, (130, 40), 1),       # 4096 + 1104
    ("zeros", 2, (128, 128), 1),        # long overlapping matches
    ("max", 2, (33, 33), 1),            # 1089 = 1088 + 1 raw
    ("poisson", 8, (32, 32), 4),        # chunks of 4 frames
    ("disk", 4, (128, 128), 1),         # ~150 sequences per block: several rounds of 64
    ("periodic", 2, (64, 64), 1),       # matches with an offset below their length
    ("noisy", 2, (64, 64), 1),          # ~40 sequences in ~2.7 KB: literals read from global memory
    ("random", 2, (64, 64), 1),         # incompressible: one 8 KB literal run from global memory
    ("runs64", 2, (64, 64), 1),         # ~500 sequences in ~1.7 KB: more than the sequence table holds
    ("runs32", 2, (64, 64), 1),         # ~1000 sequences in ~3.2 KB, the same from global memory
])
def test_decode_is_bit_exact(cuda, kind, n, shape, per_chunk):
    from quantem.gpu_amd.ssb.compute.hip.bslz4 import decode_u16
    v = _frames(kind, n, shape, n + shape[0])
    chunks = [B.encode_chunk(v[i:i + per_chunk]) for i in range(0, n, per_chunk)]
    out = decode_u16(chunks, (per_chunk,) + shape)
    np.testing.assert_array_equal(out.cpu().numpy(), v)


@pytest.mark.parametrize("block_bytes", [16, 256, 2048])
@pytest.mark.parametrize("kind", ["poisson", "runs32"])
def test_decode_other_block_sizes(cuda, kind, block_bytes):
    """Filter block sizes other than 8192 bytes: the sequence table holds
    block_bytes / 32 entries (at least 64, at most what a block can hold),
    so these blocks also reach the walk inside the decode kernel."""
    from quantem.gpu_amd.ssb.compute.hip.bslz4 import decode_u16
    v = _frames(kind, 3, (40, 64), block_bytes)
    chunks = [B.encode_chunk(v[i:i + 1], block_bytes) for i in range(3)]
    out = decode_u16(chunks, (1, 40, 64), block_bytes=block_bytes)
    np.testing.assert_array_equal(out.cpu().numpy(), v)


def test_malformed_chunks_are_rejected(cuda):
    from quantem.gpu_amd.ssb.compute.hip.bslz4 import decode_u16
    v = _frames("poisson", 1, (64, 64), 0)
    good = B.encode_chunk(v)
    with pytest.raises(ValueError, match="malformed"):
        decode_u16([good[:-3]], (1, 64, 64))                          # truncated block
    with pytest.raises(ValueError, match="malformed"):
        decode_u16([good], (1, 64, 32))                               # header disagrees
    bad = bytearray(good)
    cs = struct.unpack(">I", bad[12:16])[0]
    bad[16 + cs - 7] ^= 0xFF                                         # corrupt the stream's tail
    bad[16 + 1:16 + 3] = b"\xff\xff"                                 # an offset past the output
    with pytest.raises(ValueError, match="malformed"):
        decode_u16([bytes(bad)], (1, 64, 64))
    out = decode_u16([good], (1, 64, 64))                             # the decoder recovers
    np.testing.assert_array_equal(out.cpu().numpy(), v)


def test_decoded_dataset_reconstructs_like_the_raw_one(cuda):
    """The reference's Poisson fixture (128x128 scan, 16x16 detector), one
    frame per chunk: decode on the GPU, then the R1 reconstruction."""
    from quantem.gpu_amd.ssb.compute.hip.bslz4 import decode_u16
    data = C.poisson_data(123)
    frames = data.reshape(-1, 16, 16)
    chunks = [B.encode_chunk(f) for f in frames]
    dec = decode_u16(chunks, (1, 16, 16)).reshape(data.shape)
    raw = cuda.from_numpy(data).cuda()
    assert cuda.equal(dec, raw)
    eng = engine_for_data(cuda, data, semiangle_mrad=21.4, det_sampling=1.0, step=0.5)
    a = cuda.empty((128, 128), dtype=cuda.float32, device="cuda")
    b = cuda.empty_like(a)
    eng.recon_u16(raw, *C.MAKE_ENGINE_ABERRATIONS, 0.0, 0.0, 3.0, phase_out=a)
    eng.recon_u16(dec, *C.MAKE_ENGINE_ABERRATIONS, 0.0, 0.0, 3.0, phase_out=b)
    assert cuda.equal(a, b)


def test_reference_fixture_chunk_decodes_on_the_gpu(cuda, golden):
    """The chunk of the reference's native-IO fixture file, decoded on the
    GPU (alone and as the middle of three chunks), equals ``pixel % 257``."""
    from pathlib import Path

    from quantem.gpu_amd.ssb.compute.hip.bslz4 import decode_u16
    meta = golden["bslz4_fixture_chunk"]
    chunk = (Path(__file__).resolve().parent / "golden" / meta["file"]).read_bytes()
    expect = (np.arange(64 * 64) % 257).astype(np.uint16).reshape(1, 64, 64)
    np.testing.assert_array_equal(decode_u16([chunk], (1, 64, 64)).cpu().numpy(), expect)
    other = _frames("poisson", 2, (64, 64), 7)
    chunks = [B.encode_chunk(other[0]), chunk, B.encode_chunk(other[1])]
    out = decode_u16(chunks, (1, 64, 64)).cpu().numpy()
    np.testing.assert_array_equal(out[1], expect[0])
    np.testing.assert_array_equal(out[[0, 2]], other)
