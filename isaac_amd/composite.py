"""Sort-last compositing of sub-volume images rendered on one GPU: the
per-GPU part of the reference's IceT step (lib/isaac.hpp:800-831, 2015-2056),
through the C-ABI (include/isaac_hip.h, isaac_hip_composite_order /
isaac_hip_composite)."""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from ._lib import check, lib


def composite_order(positions: Sequence[Sequence[int]], local_sizes: Sequence[Sequence[int]],
                    global_size: Sequence[int], scale: Sequence[float], modelview) -> List[int]:
    """Sub-volume indices nearest first, as icetCompositeOrder receives them
    (lib/isaac.hpp:2021-2048).  modelview: 16 column-major doubles."""
    n = len(positions)
    pos = np.ascontiguousarray(np.asarray(positions, np.int32).reshape(n, 3))
    loc = np.ascontiguousarray(np.asarray(local_sizes, np.int32).reshape(n, 3))
    gs = np.ascontiguousarray(np.asarray(global_size, np.int32).reshape(3))
    sc = np.ascontiguousarray(np.asarray(scale, np.float32).reshape(3))
    mv = np.ascontiguousarray(np.asarray(modelview, np.float64).reshape(16))
    order = np.zeros(n, np.int32)
    check(lib().isaac_hip_composite_order(n, pos.ctypes.data, loc.ctypes.data, gs.ctypes.data, sc.ctypes.data,
                                          mv.ctypes.data, order.ctypes.data), None, "composite_order")
    return [int(v) for v in order]


def composite(images, order: Sequence[int], out=None):
    """images: CUDA uint8 tensors of equal shape (..., 4) (RGBA8, 16-byte aligned);
    returns images[order[0]] over images[order[1]] over ... (IceT's byte "over",
    pairwise tree over the order) in `out` (allocated when None), enqueued on
    torch's current stream."""
    import torch

    imgs = list(images)
    if not imgs:
        raise ValueError("no images")
    shape = imgs[0].shape
    for t in imgs:
        if not t.is_cuda or t.dtype != torch.uint8 or t.shape != shape or not t.is_contiguous():
            raise TypeError("images must be contiguous CUDA uint8 tensors of one shape")
    if out is None:
        out = torch.empty_like(imgs[0])
    n = len(imgs)
    ptrs = (ctypes.c_void_p * n)(*[t.data_ptr() for t in imgs])
    o = np.ascontiguousarray(np.asarray(order, np.int32))
    if o.shape != (n,):
        raise ValueError("order must list every image once")
    stream = torch.cuda.current_stream(imgs[0].device).cuda_stream
    check(lib().isaac_hip_composite(n, ctypes.cast(ptrs, ctypes.c_void_p), o.ctypes.data, imgs[0].numel() // 4,
                                    ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)), None, "composite")
    return out


__all__ = ["composite_order", "composite"]
