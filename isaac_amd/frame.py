"""The reference's video frame message built from a frame in device memory
(serverSendFrame without libjpeg, lib/isaac/isaac_communicator.hpp:217-287),
through the C-ABI (isaac_hip_base64, isaac_hip_frame_message)."""
from __future__ import annotations

import ctypes

from ._lib import check, lib


def base64_length(n: int) -> int:
    return (8 * n + 5) // 6


def base64(data, out=None):
    """Device bytes (a contiguous CUDA uint8 tensor) -> device base64 characters
    (uint8 tensor of ceil(8n/6)), no '=' padding; enqueued on torch's current stream."""
    import torch

    if not data.is_cuda or data.dtype != torch.uint8 or not data.is_contiguous():
        raise TypeError("data must be a contiguous CUDA uint8 tensor")
    n = data.numel()
    if out is None:
        out = torch.empty(base64_length(n), dtype=torch.uint8, device=data.device)
    stream = torch.cuda.current_stream(data.device).cuda_stream
    check(lib().isaac_hip_base64(ctypes.c_void_p(data.data_ptr()), n, ctypes.c_void_p(out.data_ptr()),
                                 ctypes.c_void_p(stream)), None, "base64")
    return out


def frame_message(rgba) -> bytes:
    """{"payload": "data:image/raw-rgba;base64,..."} for an (H, W, 4) CUDA uint8 frame."""
    import torch

    if not rgba.is_cuda or rgba.dtype != torch.uint8 or not rgba.is_contiguous() or rgba.dim() != 3 or rgba.shape[2] != 4:
        raise TypeError("rgba must be a contiguous CUDA uint8 tensor of shape (H, W, 4)")
    H, W = int(rgba.shape[0]), int(rgba.shape[1])
    scratch = torch.empty(max(1, base64_length(4 * W * H)), dtype=torch.uint8, device=rgba.device)
    cap = base64_length(4 * W * H) + 64
    buf = ctypes.create_string_buffer(cap)
    length = ctypes.c_int64(0)
    stream = torch.cuda.current_stream(rgba.device).cuda_stream
    check(lib().isaac_hip_frame_message(ctypes.c_void_p(rgba.data_ptr()), W, H, ctypes.c_void_p(scratch.data_ptr()), buf,
                                        cap, ctypes.byref(length), ctypes.c_void_p(stream)), None, "frame_message")
    return buf.raw[:length.value]


__all__ = ["base64", "base64_length", "frame_message"]
