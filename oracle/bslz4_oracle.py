"""CPU restatement of HDF5 bitshuffle-LZ4 ("bslz4") detector chunks.

TEST INFRASTRUCTURE ONLY: ``tests/`` use it to write compressed fixtures and
to check the GPU decoder of ``quantem/gpu_amd/csrc/ssb_bslz4.hip``; nothing
in the package imports it.

Format, as the reference's readers consume it (``src/quantem/gpu/io/backends/
webgpu/bslz4.ts:1-30``, ``140-167``; ``src/quantem/gpu/io/backends/cuda/decoder.py:954-1000``;
written with hdf5plugin by ``tests/io/test_mps_buffer_release.py:241-330``):

* chunk = big-endian uint64 uncompressed bytes, big-endian uint32 block
  bytes, then per block a big-endian uint32 compressed size and an LZ4
  block, then the raw tail;
* a block holds ``block_bytes / elem_size`` elements; after the full blocks
  a last block holds the remaining elements rounded down to a multiple of 8,
  and the final ``n mod 8`` elements are stored raw (little-endian);
* within a block of ``m`` elements of ``s`` bytes the data are bit-shuffled
  into ``8 s`` bit-planes of ``m / 8`` bytes: bit ``b`` of element ``p`` is
  bit ``p mod 8`` of byte ``b * m/8 + p // 8`` (plane-major, LSB first);
* an LZ4 block (the public LZ4 block format) is a run of sequences: a token
  whose high nibble is the literal count and low nibble the match length
  minus 4 (15 = more bytes follow, each adding up to 255), the literals, a
  little-endian uint16 offset and the extra match-length bytes; the last
  sequence is literals only.

The hdf5plugin/bitshuffle library is not available here, so the byte
layout is pinned by these published descriptions and by hand-assembled
known answers in ``tests/test_bslz4_oracle.py``, not by files it wrote.
"""
from __future__ import annotations

import struct

import numpy as np

MIN_MATCH = 4
LAST_LITERALS = 5        # the last 5 bytes of a block are always literals
MFLIMIT = 12             # a match may not start in the last 12 bytes
DEFAULT_BLOCK_BYTES = 8192


def bitshuffle(block: np.ndarray) -> bytes:
    """Bit-planes of ``m`` elements (``m`` a multiple of 8), plane-major, LSB first."""
    a = np.ascontiguousarray(block)
    m, s = a.size, a.dtype.itemsize
    if m % 8:
        raise ValueError("bitshuffle blocks hold a multiple of 8 elements")
    bits = np.unpackbits(a.view(np.uint8).reshape(m, s), axis=1, bitorder="little")   # [p][b]
    return np.packbits(bits.T, axis=1, bitorder="little").tobytes()                    # [b][p/8]


def bitunshuffle(data: bytes, m: int, dtype) -> np.ndarray:
    dt = np.dtype(dtype)
    s = dt.itemsize
    planes = np.frombuffer(data, dtype=np.uint8, count=m * s).reshape(8 * s, m // 8)
    bits = np.unpackbits(planes, axis=1, bitorder="little")                            # [b][p]
    return np.packbits(bits.T, axis=1, bitorder="little").reshape(-1).view(dt).copy()


def lz4_compress(src: bytes) -> bytes:
    """Greedy LZ4 block compressor (4-byte hash, one candidate per hash)."""
    n = len(src)
    out = bytearray()
    table: dict[bytes, int] = {}
    anchor = i = 0
    limit = n - MFLIMIT

    def put_len(x: int) -> None:
        while x >= 255:
            out.append(255)
            x -= 255
        out.append(x)

    while i <= limit:
        key = src[i:i + 4]
        cand = table.get(key)
        table[key] = i
        if cand is None or i - cand > 65535:
            i += 1
            continue
        length = 4
        while i + length < n - LAST_LITERALS and src[cand + length] == src[i + length]:
            length += 1
        lit = i - anchor
        ml = length - MIN_MATCH
        out.append((min(lit, 15) << 4) | min(ml, 15))
        if lit >= 15:
            put_len(lit - 15)
        out += src[anchor:i]
        out += struct.pack("<H", i - cand)
        if ml >= 15:
            put_len(ml - 15)
        i += length
        anchor = i
    lit = n - anchor
    out.append(min(lit, 15) << 4)
    if lit >= 15:
        put_len(lit - 15)
    out += src[anchor:]
    return bytes(out)


def lz4_decompress(src: bytes, out_size: int) -> bytes:
    """LZ4 block decoder; raises ValueError on any out-of-range read or write."""
    out = bytearray()
    i = 0
    n = len(src)

    def get_len(base: int) -> int:
        nonlocal i
        if base != 15:
            return base
        while True:
            if i >= n:
                raise ValueError("truncated length")
            b = src[i]
            i += 1
            base += b
            if b != 255:
                return base

    while True:
        if i >= n:
            raise ValueError("truncated sequence")
        token = src[i]
        i += 1
        lit = get_len(token >> 4)
        if i + lit > n or len(out) + lit > out_size:
            raise ValueError("literals out of range")
        out += src[i:i + lit]
        i += lit
        if i == n:
            break
        if i + 2 > n:
            raise ValueError("truncated offset")
        off = src[i] | (src[i + 1] << 8)
        i += 2
        ml = get_len(token & 15) + MIN_MATCH
        if off == 0 or off > len(out) or len(out) + ml > out_size:
            raise ValueError("match out of range")
        start = len(out) - off
        for k in range(ml):                      # byte by byte: overlapping copies repeat
            out.append(out[start + k])
    if len(out) != out_size:
        raise ValueError("decoded size mismatch")
    return bytes(out)


def block_plan(n_elems: int, block_elems: int) -> tuple[int, int, int]:
    """(full blocks, elements of the last shuffled block, raw tail elements)."""
    full = n_elems // block_elems
    rest = n_elems - full * block_elems
    return full, rest - rest % 8, rest % 8


def encode_chunk(values: np.ndarray, block_bytes: int = DEFAULT_BLOCK_BYTES) -> bytes:
    a = np.ascontiguousarray(values).reshape(-1)
    s = a.dtype.itemsize
    be = block_bytes // s
    full, last, tail = block_plan(a.size, be)
    out = bytearray(struct.pack(">QI", a.size * s, block_bytes))
    pos = 0
    for m in [be] * full + ([last] if last else []):
        comp = lz4_compress(bitshuffle(a[pos:pos + m]))
        out += struct.pack(">I", len(comp)) + comp
        pos += m
    out += a[pos:].astype(a.dtype.newbyteorder("<")).tobytes()
    return bytes(out)


def decode_chunk(chunk: bytes, dtype=np.uint16) -> np.ndarray:
    dt = np.dtype(dtype)
    s = dt.itemsize
    total, block_bytes = struct.unpack(">QI", chunk[:12])
    n = total // s
    be = block_bytes // s
    full, last, tail = block_plan(n, be)
    parts = []
    p = 12
    for m in [be] * full + ([last] if last else []):
        (cs,) = struct.unpack(">I", chunk[p:p + 4])
        p += 4
        parts.append(bitunshuffle(lz4_decompress(chunk[p:p + cs], m * s), m, dt))
        p += cs
    parts.append(np.frombuffer(chunk[p:p + tail * s], dtype=dt.newbyteorder("<")).astype(dt))
    if p + tail * s != len(chunk):
        raise ValueError("chunk length does not match its blocks")
    return np.concatenate(parts)
