"""JPEG frames on the GPU (isaac_hip_jpeg_encode) byte-identical to
libjpeg-turbo (Pillow's bundled copy, the library the reference links) on
frames that exercise every edge rule, at 1080p and on a rendered frame, and to
the oracle for qualities below 25 (16-bit tables, which Pillow does not
write); the frame message equals the oracle's."""
import io

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu
PIL = pytest.importorskip("PIL.Image")


def _pil(rgb, q):
    b = io.BytesIO()
    PIL.fromarray(rgb).save(b, "JPEG", quality=q, subsampling=2, optimize=False, progressive=False)
    return b.getvalue()


def _frame(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    rgb = np.stack([(xx * 7 + yy * 3) % 256, (xx * 2 + yy * 11) % 256, (xx * yy) % 256, np.full_like(xx, 255)], -1)
    return np.clip(rgb + rng.integers(-25, 25, size=rgb.shape), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("shape", [(1, 1), (8, 8), (17, 23), (33, 50), (47, 81), (120, 200), (1080, 1920)])
@pytest.mark.parametrize("quality", [30, 90, 100])
def test_byte_identical_to_libjpeg_turbo(gpu, shape, quality):
    import torch

    from isaac_amd import frame as K

    rgba = _frame(*shape, seed=shape[0] + 7 * shape[1])
    got = K.jpeg(torch.from_numpy(rgba).to(gpu), quality)
    want = _pil(np.ascontiguousarray(rgba[..., :3]), quality)
    print(f"{shape} q{quality}: {len(got)} bytes (libjpeg-turbo {len(want)})", flush=True)
    assert got == want


@pytest.mark.parametrize("quality", [1, 10, 24])
def test_low_quality_equals_oracle(gpu, oracle_lib, quality):
    import torch

    from isaac_amd import frame as K
    from oracle import jpeg_oracle as J

    rgba = _frame(33, 50, 3)
    assert K.jpeg(torch.from_numpy(rgba).to(gpu), quality) == J.encode(rgba, quality)


def test_rendered_frame_and_message(gpu, oracle_lib):
    import torch

    from isaac_amd import frame as K
    from isaac_amd import visualization as V
    from oracle import jpeg_oracle as J

    spec = scenes.small_scene(n=32, fb=(333, 201))
    vis = V.from_scene(spec, gpu.index or 0)
    vis.update_buffers()
    img = vis.draw()
    vis.close()
    dev = torch.from_numpy(img).to(gpu)
    assert K.jpeg(dev, 90) == _pil(np.ascontiguousarray(img[..., :3]), 90)
    assert K.jpeg_frame_message(dev, 90) == J.frame_message(img, 90)


def test_video_frame_message_of_the_context(gpu, oracle_lib):
    from isaac_amd import visualization as V
    from isaac_amd._lib import IsaacHipError
    from oracle import frame_oracle as F
    from oracle import jpeg_oracle as J

    spec = scenes.small_scene(n=32, fb=(250, 130))
    vis = V.from_scene(spec, gpu.index or 0)
    with pytest.raises(IsaacHipError):
        vis.video_frame_message(90)  # nothing rendered yet
    vis.update_buffers()
    img = vis.draw()
    assert vis.video_frame_message(90) == J.frame_message(img, 90)
    assert vis.video_frame_message(0) == F.frame_message(img.tobytes())
    vis.close()
