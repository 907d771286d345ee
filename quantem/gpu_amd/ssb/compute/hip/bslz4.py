"""GPU decode of HDF5 bitshuffle-LZ4 detector chunks into the resident
uint16 frame array (``ssb_bslz4_decode_u16``).

The caller hands over the raw chunk bytes as stored in the HDF5 file (what
``h5py``'s ``read_direct_chunk`` returns for each ``(frames, rows, cols)``
chunk of an Arina/Dectris master), concatenated; the decode runs on the
GPU, so only compressed bytes cross PCIe.  Replaces the reference's CUDA
decoder (``src/quantem/gpu/io/backends/cuda/decoder.py:258-1146``).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi

DEFAULT_BLOCK_BYTES = 8192


class Bslz4Chunks:
    """Compressed chunks resident on the GPU, decodable repeatedly.

    ``chunks``: a list of ``bytes`` (one per chunk), or a pair
    ``(buffer, sizes)`` of one uint8 buffer (host array or GPU tensor) and
    the per-chunk byte counts.  ``chunk_shape``: the decoded shape of one
    chunk, e.g. ``(1, rows, cols)`` for one frame per chunk."""

    def __init__(self, chunks, chunk_shape, block_bytes: int = DEFAULT_BLOCK_BYTES, device=None):
        import torch
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if isinstance(chunks, (list, tuple)) and (not chunks or isinstance(chunks[0], (bytes, bytearray))):
            sizes = np.array([len(c) for c in chunks], dtype=np.int64)
            buf = np.frombuffer(b"".join(chunks), dtype=np.uint8)
        else:
            buf, sizes = chunks
            sizes = np.asarray(sizes, dtype=np.int64)
        self.n = int(sizes.size)
        offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64) if self.n else np.zeros(0, np.int64)
        comp = buf if isinstance(buf, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(buf))
        self.comp = comp.to(self.device, dtype=torch.uint8).contiguous()
        if self.comp.numel() < int(sizes.sum()):
            raise ValueError("chunk sizes exceed the buffer")
        self.compressed_bytes = int(sizes.sum())
        self.chunk_shape = tuple(int(v) for v in chunk_shape)
        self.chunk_elems = int(np.prod(self.chunk_shape))
        self.block_bytes = int(block_bytes)
        self.shape = (self.n * self.chunk_shape[0],) + self.chunk_shape[1:]
        self.offsets = torch.from_numpy(offsets).to(self.device)
        self.sizes = torch.from_numpy(sizes).to(self.device)
        need = int(_abi.lib().ssb_bslz4_workspace_bytes(self.n, self.chunk_elems, self.block_bytes))
        if need < 0:
            _abi.check(need, "ssb_bslz4_workspace_bytes")
        self.workspace = torch.empty(max(need, 256), dtype=torch.uint8, device=self.device)

    def decode(self, out=None, sync: bool = True):
        """Decode into ``out`` (new uint16 tensor by default).  With ``sync`` a
        malformed chunk raises ``ValueError``; otherwise call ``check``."""
        import torch
        if out is None:
            out = torch.empty(self.shape, dtype=torch.uint16, device=self.device)
        elif out.dtype != torch.uint16 or out.numel() != self.n * self.chunk_elems or not out.is_contiguous():
            raise ValueError("out must be a contiguous uint16 tensor of the decoded size")
        elif not out.is_cuda or out.device != self.device:
            # the decode kernel writes through out's address
            raise ValueError(f"out must live on {self.device}; got {out.device}")
        status = ctypes.c_int(0)
        _abi.call("ssb_bslz4_decode_u16", _abi.ptr(self.comp), int(self.comp.numel()), _abi.ptr(self.offsets),
                  _abi.ptr(self.sizes), self.n, self.chunk_elems, self.block_bytes, _abi.ptr(out),
                  _abi.ptr(self.workspace), int(self.workspace.numel()), ctypes.byref(status) if sync else None,
                  _abi.stream_handle(torch.cuda.current_stream(self.device)))
        return out

    def check(self) -> None:
        """Raise ``ValueError`` if the last asynchronous decode met a malformed chunk."""
        code = int(self.workspace[:4].view(dtype=__import__("torch").int32).item())
        if code:
            raise ValueError(f"malformed bitshuffle-LZ4 chunk (status {code})")


def decode_u16(chunks, chunk_shape, block_bytes: int = DEFAULT_BLOCK_BYTES, device=None, out=None):
    """Decode bslz4 chunks to a uint16 tensor of shape
    ``(n_chunks * chunk_shape[0], *chunk_shape[1:])`` on the GPU; a malformed
    chunk raises ``ValueError``."""
    return Bslz4Chunks(chunks, chunk_shape, block_bytes, device).decode(out=out)


def encode_u16(values, block_bytes: int = DEFAULT_BLOCK_BYTES) -> bytes:
    """One bitshuffle-LZ4 chunk of host uint16 ``values`` (host-side encoder
    of the native library; the layout the decoder reads)."""
    v = np.ascontiguousarray(np.asarray(values, dtype=np.uint16).reshape(-1))
    lib = _abi.lib()
    bound = int(lib.ssb_bslz4_encode_bound(int(v.size), int(block_bytes)))
    if bound < 0:
        _abi.check(bound, "ssb_bslz4_encode_bound")
    out = np.empty(bound, dtype=np.uint8)
    n = int(lib.ssb_bslz4_encode_u16(v.ctypes.data_as(ctypes.c_void_p), int(v.size), int(block_bytes),
                                     out.ctypes.data_as(ctypes.c_void_p), bound))
    if n < 0:
        _abi.check(n, "ssb_bslz4_encode_u16")
    return out[:n].tobytes()
