"""The bslz4 oracle pinned by known answers: the reference's documented
bit-plane order, hand-assembled LZ4 blocks of the public block format,
and encode/decode round trips over the edge cases of the chunk layout."""
from __future__ import annotations

import struct

import numpy as np
import pytest

from oracle import bslz4_oracle as B


def test_bit_plane_order_is_plane_major_lsb_first():
    # bit b of element p is bit (p mod 8) of byte b*m/8 + p//8 (bslz4.ts:140-160)
    v = np.zeros(16, dtype=np.uint16)
    v[0] = 1            # bit 0 of element 0 -> plane 0, byte 0, bit 0
    v[3] = 0x8000       # bit 15 of element 3 -> plane 15, byte 0, bit 3
    v[9] = 0x0100       # bit 8 of element 9 -> plane 8, byte 1, bit 1
    out = np.frombuffer(B.bitshuffle(v), dtype=np.uint8).reshape(16, 2)
    expect = np.zeros((16, 2), dtype=np.uint8)
    expect[0, 0] = 0x01
    expect[15, 0] = 0x08
    expect[8, 1] = 0x02
    np.testing.assert_array_equal(out, expect)
    const = np.full(8, 0xA5A5, dtype=np.uint16)     # constant block: planes are 0x00 or 0xFF
    planes = np.frombuffer(B.bitshuffle(const), dtype=np.uint8)
    assert set(planes.tolist()) == {0x00, 0xFF}
    np.testing.assert_array_equal(B.bitunshuffle(B.bitshuffle(v), 16, np.uint16), v)


def test_lz4_known_blocks():
    assert B.lz4_decompress(bytes([0x30]) + b"abc", 3) == b"abc"
    # 'a', then a match of offset 1 and length 19 (overlapping copy), then 5 literals
    blk = bytes([0x1F]) + b"a" + struct.pack("<H", 1) + bytes([0]) + bytes([0x50]) + b"bcdef"
    assert B.lz4_decompress(blk, 25) == b"a" * 20 + b"bcdef"
    # literal length 20 = 15 + 5 (one extra byte), offset 4, match 4 + 15 + 255 + 1
    lits = bytes(range(20))
    blk = bytes([0xFF, 5]) + lits + struct.pack("<H", 4) + bytes([255, 1]) + bytes([0x50]) + b"vwxyz"
    out = B.lz4_decompress(blk, 20 + 275 + 5)
    assert out[:20] == lits and out[20:295] == (lits[16:20] * 69)[:275] and out[295:] == b"vwxyz"
    with pytest.raises(ValueError):
        B.lz4_decompress(bytes([0x1F]) + b"a" + struct.pack("<H", 2) + bytes([0]), 20)   # offset too far
    with pytest.raises(ValueError):
        B.lz4_decompress(bytes([0x30]) + b"ab", 3)                                      # truncated


@pytest.mark.parametrize("n,kind", [(16384, "poisson"), (4096 * 3 + 64 + 5, "poisson"), (7, "random"),
                                    (9000, "random"), (20000, "zeros"), (8192, "max")])
def test_chunk_round_trip_and_layout(n, kind):
    rng = np.random.default_rng(n)
    v = {"poisson": lambda: rng.poisson(0.7, n), "random": lambda: rng.integers(0, 65536, n),
         "zeros": lambda: np.zeros(n), "max": lambda: np.full(n, 65535)}[kind]().astype(np.uint16)
    chunk = B.encode_chunk(v)
    total, bb = struct.unpack(">QI", chunk[:12])
    assert total == 2 * n and bb == 8192
    full, last, tail = B.block_plan(n, 4096)
    assert full * 4096 + last + tail == n and last % 8 == 0 and tail < 8
    np.testing.assert_array_equal(B.decode_chunk(chunk), v)
    if kind in ("zeros", "max"):
        assert len(chunk) < n // 20              # long runs compress


@pytest.mark.parametrize("n,kind", [(16384, "poisson"), (9000, "random"), (49, "random"), (20000, "zeros")])
def test_library_encoder_output_decodes_with_the_oracle(n, kind):
    """The native host encoder writes chunks the oracle decodes exactly (no GPU)."""
    from quantem.gpu_amd.ssb.compute.hip.bslz4 import encode_u16
    rng = np.random.default_rng(n + 1)
    v = {"poisson": lambda: rng.poisson(0.7, n), "random": lambda: rng.integers(0, 65536, n),
         "zeros": lambda: np.zeros(n)}[kind]().astype(np.uint16)
    chunk = encode_u16(v)
    np.testing.assert_array_equal(B.decode_chunk(chunk), v)
    assert chunk[:12] == B.encode_chunk(v)[:12]
    if kind == "zeros":
        assert len(chunk) < n // 20


def test_reference_fixture_chunk_decodes_to_its_known_frame(golden):
    """The chunk of the reference's native-IO fixture file decodes to
    ``pixel % 257`` on 64 x 64 (``Native4DSTEMIOTests.swift:1224-1226``)."""
    from pathlib import Path
    meta = golden["bslz4_fixture_chunk"]
    chunk = (Path(__file__).resolve().parent / "golden" / meta["file"]).read_bytes()
    assert len(chunk) == meta["bytes"] == 778
    out = B.decode_chunk(chunk)
    np.testing.assert_array_equal(out, (np.arange(64 * 64) % 257).astype(np.uint16))
    # a chunk written by the reference's encoder, not by ours: the oracle's
    # encoder need not produce the same bytes, only the same frame
    assert np.array_equal(B.decode_chunk(B.encode_chunk(out)), out)
