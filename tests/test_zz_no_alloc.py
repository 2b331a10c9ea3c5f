"""isaac_hip_jpeg_encode_noalloc, the JPEG encoder that allocates nothing, against
isaac_hip_jpeg_encode with a caller workspace: byte-identical files on two calls in a
row (a 1920x1080 frame, then a smaller one, so the second call runs in scratch that
still holds the first call's contents), frames too large for the scratch rejected
before any work, and a stream of another device rejected.  Every call goes through
the C-ABI."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _frame(h, w, seed):
    """A smooth gradient with noise in half of it: both short and long Huffman runs."""
    import torch

    g = torch.Generator(device="cpu").manual_seed(seed)
    y = torch.arange(h, dtype=torch.float32)[:, None].expand(h, w)
    x = torch.arange(w, dtype=torch.float32)[None, :].expand(h, w)
    base = torch.stack([x / w * 255.0, y / h * 255.0, (x + y) / (w + h) * 255.0], dim=-1)
    noise = torch.randint(0, 256, (h, w, 3), generator=g).float()
    mask = (x > w / 2).float()[..., None]
    rgb = base * (1 - mask) + noise * mask
    a = torch.full((h, w, 1), 255.0)
    return torch.cat([rgb, a], dim=-1).round().to(torch.uint8).contiguous()


def _with_workspace(dev_frame, q):
    import torch

    from isaac_amd._lib import lib

    h, w = int(dev_frame.shape[0]), int(dev_frame.shape[1])
    nbytes = int(lib().isaac_hip_jpeg_workspace_bytes(w, h))
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev_frame.device)
    buf = ctypes.create_string_buffer(2 * nbytes + 8192)
    n = ctypes.c_int64(0)
    stream = torch.cuda.current_stream(dev_frame.device).cuda_stream
    rc = lib().isaac_hip_jpeg_encode(ctypes.c_void_p(dev_frame.data_ptr()), w, h, q, ctypes.c_void_p(ws.data_ptr()), nbytes,
                                     buf, len(buf), ctypes.byref(n), ctypes.c_void_p(stream))
    assert rc == 0
    return buf.raw[:n.value]


def _noalloc(dev_frame, q, stream=None):
    import torch

    from isaac_amd._lib import lib

    h, w = int(dev_frame.shape[0]), int(dev_frame.shape[1])
    cap = 2 * int(lib().isaac_hip_jpeg_workspace_bytes(w, h)) + 8192
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_int64(0)
    if stream is None:
        stream = torch.cuda.current_stream(dev_frame.device).cuda_stream
    rc = lib().isaac_hip_jpeg_encode_noalloc(ctypes.c_void_p(dev_frame.data_ptr()), w, h, q, buf, cap, ctypes.byref(n),
                                             ctypes.c_void_p(stream))
    return rc, buf.raw[:n.value]


def test_noalloc_jpeg_equals_workspace_encoder(gpu):
    """Two calls in a row, the second on a smaller frame, each byte-identical to the
    workspace encoder's file; then the first frame again."""
    frames = [(_frame(1080, 1920, 1), 90), (_frame(301, 517, 2), 75), (_frame(1080, 1920, 3), 90)]
    for host, q in frames:
        dev = host.to(gpu)
        want = _with_workspace(dev, q)
        rc, got = _noalloc(dev, q)
        assert rc == 0
        assert got[:2] == b"\xff\xd8" and got[-2:] == b"\xff\xd9"
        print(f"{tuple(host.shape[:2])} q{q}: {len(got)} bytes", flush=True)
        assert got == want, (tuple(host.shape), q)


def test_noalloc_jpeg_limits(gpu):
    """Frames whose workspace exceeds the scratch, bad sizes and a misaligned frame
    are refused on the host, before anything is enqueued."""
    import torch

    from isaac_amd._lib import lib

    L = lib()
    cap = 1 << 20
    buf = ctypes.create_string_buffer(cap)
    n = ctypes.c_int64(0)
    dev = torch.zeros((8, 8, 4), dtype=torch.uint8, device=gpu)
    assert L.isaac_hip_jpeg_workspace_bytes(2560, 1440) > (44 << 20)
    assert L.isaac_hip_jpeg_workspace_bytes(1920, 1080) <= (44 << 20)
    # sizes are checked before the frame is read: the 8x8 tensor is never touched
    assert L.isaac_hip_jpeg_encode_noalloc(ctypes.c_void_p(dev.data_ptr()), 2560, 1440, 90, buf, cap, ctypes.byref(n), None) == -1
    assert L.isaac_hip_jpeg_encode_noalloc(ctypes.c_void_p(dev.data_ptr()), 0, 8, 90, buf, cap, ctypes.byref(n), None) == -1
    assert L.isaac_hip_jpeg_encode_noalloc(ctypes.c_void_p(dev.data_ptr() + 1), 8, 8, 90, buf, cap, ctypes.byref(n), None) == -1
    # a capacity too small reports the length and fails; the file then fits
    rc = L.isaac_hip_jpeg_encode_noalloc(ctypes.c_void_p(dev.data_ptr()), 8, 8, 90, buf, 16, ctypes.byref(n), None)
    assert rc == -1 and n.value > 16
    rc, got = _noalloc(dev, 90)
    assert rc == 0 and got == _with_workspace(dev, 90)


def test_noalloc_jpeg_on_torch_stream(gpu):
    """On a side stream of the current device, ordered after the frame's producer."""
    import torch

    host = _frame(540, 960, 4)
    s = torch.cuda.Stream(device=gpu)
    with torch.cuda.stream(s):
        dev = host.to(gpu, non_blocking=False)
        rc, got = _noalloc(dev, 85, stream=s.cuda_stream)
    assert rc == 0
    assert got == _with_workspace(dev, 85)
    assert np.frombuffer(got, np.uint8).size > 1000
