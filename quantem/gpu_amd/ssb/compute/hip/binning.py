"""Integer detector binning of resident uint16 frames on the GPU
(``ssb_bin_u16``).

``bin_frames(frames, det_bin)`` sums ``det_bin x det_bin`` detector blocks
exactly, with dead pixels (``pixel_mask`` nonzero) zeroed first and the
trailing rows/columns that do not fill a block trimmed:

* ``dtype=torch.uint32`` (default): the exact sums, the widened output of
  the reference's binned loads (``src/quantem/gpu/io/load.py:3009-3026``,
  ``3398-3415``);
* ``dtype=torch.uint16``: the low 16 bits, as the reference's CPU loader's
  uint64 sum cast back to the source dtype (``_bin_sum``,
  ``src/quantem/gpu/io/backends/cpu/reference.py:27-42``; mask at ``:96-111``).

Frames decoded by ``bslz4.Bslz4Chunks.decode`` can be binned in place on
the device; nothing crosses PCIe.
"""
from __future__ import annotations

from . import _abi


def bin_frames(frames, det_bin: int, pixel_mask=None, dtype=None, out=None):
    """Bin the last two axes of a uint16 CUDA tensor ``(..., rows, cols)``.

    ``pixel_mask``: ``(rows, cols)`` array or tensor, nonzero = dead pixel.
    Returns a new ``(..., rows // det_bin, cols // det_bin)`` tensor of
    ``dtype`` (uint32 by default), or fills ``out``."""
    import torch
    if not isinstance(frames, torch.Tensor) or frames.device.type != "cuda":
        raise TypeError("bin_frames needs a CUDA tensor (the hip library bins resident frames)")
    if frames.dtype != torch.uint16:
        raise NotImplementedError(f"bin_frames reads uint16 detector counts; got {frames.dtype}")
    if frames.ndim < 2:
        raise ValueError("frames must have detector axes (..., rows, cols)")
    dtype = torch.uint32 if dtype is None else dtype
    if dtype not in (torch.uint32, torch.uint16):
        raise ValueError("binned dtype must be torch.uint32 or torch.uint16")
    det_bin = int(det_bin)
    rows, cols = int(frames.shape[-2]), int(frames.shape[-1])
    if det_bin < 1 or det_bin > min(rows, cols, 256):
        raise ValueError(f"det_bin must be in [1, min(rows, cols, 256)]; got {det_bin}")
    src = frames.contiguous()
    n = src.numel() // (rows * cols)
    shape = tuple(src.shape[:-2]) + (rows // det_bin, cols // det_bin)
    if out is None:
        out = torch.empty(shape, dtype=dtype, device=src.device)
    elif tuple(out.shape) != shape or out.dtype != dtype or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous {dtype} tensor of shape {shape}")
    elif not isinstance(out, torch.Tensor) or not out.is_cuda or out.device != src.device:
        # the kernel writes through out's address: a host or other-GPU
        # tensor must be refused before any launch
        raise ValueError(f"out must live on {src.device}; got {out.device}")
    bad = None
    if pixel_mask is not None:
        bad = torch.as_tensor(pixel_mask).to(device=src.device)
        if tuple(bad.shape) != (rows, cols):
            raise ValueError(f"pixel_mask must have shape {(rows, cols)}; got {tuple(bad.shape)}")
        bad = (bad != 0).to(torch.uint8).contiguous()
    stream = torch.cuda.current_stream(src.device).cuda_stream
    # asynchronous on torch's current stream: the temporaries (src copy, mask)
    # are freed to the caching allocator, which reuses them only for later
    # work on this same stream
    _abi.call("ssb_bin_u16", _abi.ptr(src), n, rows, cols, det_bin, _abi.ptr(bad), 32 if dtype == torch.uint32 else 16,
              _abi.ptr(out), stream)
    return out
