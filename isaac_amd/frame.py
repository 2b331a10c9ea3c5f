"""The reference's video frame message built from a frame in device memory
(serverSendFrame, lib/isaac/isaac_communicator.hpp:217-287): the raw-RGBA
payload (isaac_hip_base64, isaac_hip_frame_message) and the default JPEG path
(isaac_hip_jpeg_encode, isaac_hip_jpeg_frame_message), through the C-ABI."""
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


_SCRATCH = {}


def _frame_args(rgba):
    """frame size, the encoder's device workspace and a host buffer for the file
    (both kept per device and frame size), and torch's current stream"""
    import torch

    if not rgba.is_cuda or rgba.dtype != torch.uint8 or not rgba.is_contiguous() or rgba.dim() != 3 or rgba.shape[2] != 4:
        raise TypeError("rgba must be a contiguous CUDA uint8 tensor of shape (H, W, 4)")
    H, W = int(rgba.shape[0]), int(rgba.shape[1])
    key = (rgba.device.index, W, H)
    if key not in _SCRATCH:
        nbytes = int(lib().isaac_hip_jpeg_workspace_bytes(W, H))
        if nbytes < 0:
            raise ValueError("JPEG frames are 1..65535 pixels a side")
        ws = torch.empty(nbytes, dtype=torch.uint8, device=rgba.device)
        _SCRATCH[key] = (ws, ctypes.create_string_buffer(2 * nbytes + 8192))
    ws, host = _SCRATCH[key]
    stream = torch.cuda.current_stream(rgba.device).cuda_stream
    return W, H, ws, host, stream


def jpeg(rgba, quality: int = 90) -> bytes:
    """The JPEG file libjpeg-turbo writes for the frame (JCS_EXT_RGBX, jpeg_set_defaults,
    jpeg_set_quality(quality, FALSE)), encoded on the GPU; rgba: (H, W, 4) CUDA uint8."""
    W, H, ws, host, stream = _frame_args(rgba)
    n = ctypes.c_int64(0)
    check(lib().isaac_hip_jpeg_encode(ctypes.c_void_p(rgba.data_ptr()), W, H, int(quality), ctypes.c_void_p(ws.data_ptr()),
                                      ws.numel(), host, len(host), ctypes.byref(n), ctypes.c_void_p(stream)), None,
          "jpeg_encode")
    return ctypes.string_at(host, n.value)


def jpeg_frame_message(rgba, quality: int = 90) -> bytes:
    """{"payload": "data:image/jpeg;base64,..."} for an (H, W, 4) CUDA uint8 frame."""
    W, H, ws, host, stream = _frame_args(rgba)
    n = ctypes.c_int64(0)
    check(lib().isaac_hip_jpeg_frame_message(ctypes.c_void_p(rgba.data_ptr()), W, H, int(quality),
                                             ctypes.c_void_p(ws.data_ptr()), ws.numel(), host, len(host), ctypes.byref(n),
                                             ctypes.c_void_p(stream)), None, "jpeg_frame_message")
    return ctypes.string_at(host, n.value)


__all__ = ["base64", "base64_length", "frame_message", "jpeg", "jpeg_frame_message"]
