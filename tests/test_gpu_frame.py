"""The frame message on the GPU (isaac_hip_base64, isaac_hip_frame_message)
equals the oracle's restatement byte for byte: byte counts around the 12-byte
groups and the tail, a 1080p frame, and a rendered frame."""
import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 8, 9, 10, 11, 12, 13, 23, 24, 25, 4096, 1000003])
def test_base64_bit_exact(gpu, n):
    import torch

    from isaac_amd import frame as K
    from oracle import frame_oracle as F

    rng = np.random.default_rng(n)
    data = rng.integers(0, 256, size=n, dtype=np.uint8)
    out = K.base64(torch.from_numpy(data).to(gpu))
    torch.cuda.synchronize()
    got = out.cpu().numpy().tobytes()
    want = F.b64_unpadded(data.tobytes())
    print(f"n={n}: {len(got)} characters", flush=True)
    assert got == want


def test_frame_message_of_a_rendered_frame(gpu):
    import torch

    from isaac_amd import frame as K
    from isaac_amd import visualization as V
    from oracle import frame_oracle as F

    spec = scenes.small_scene(n=32, fb=(1920, 1080))
    vis = V.from_scene(spec, gpu.index or 0)
    vis.update_buffers()
    img = vis.draw()
    vis.close()
    msg = K.frame_message(torch.from_numpy(img).to(gpu))
    want = F.frame_message(img.tobytes())
    assert msg == want
    assert len(msg) == len(b'{"payload": "data:image/raw-rgba;base64,"}') + (8 * img.size + 5) // 6
