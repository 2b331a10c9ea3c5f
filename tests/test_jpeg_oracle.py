"""The JPEG restatement (oracle/jpeg_oracle.py) against libjpeg-turbo itself:
Pillow's bundled libjpeg-turbo writes byte-identical files for the same frame
(baseline, 4:2:0, jpeg_set_defaults, quality q) on sizes that exercise every
edge rule (1x1, odd widths and heights, partial MCUs, dummy blocks) and on
qualities 30..100 (Pillow forces 8-bit tables, which only differ below 25).
Also: the integer DCT of a constant block is its DC alone, and the frame
message framing."""
import io

import numpy as np
import pytest

from oracle import jpeg_oracle as J

PIL = pytest.importorskip("PIL.Image")


def _pil(rgb, q):
    b = io.BytesIO()
    PIL.fromarray(rgb).save(b, "JPEG", quality=q, subsampling=2, optimize=False, progressive=False)
    return b.getvalue()


def _frame(h, w, seed):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    rgb = np.stack([(xx * 7 + yy * 3) % 256, (xx * 2 + yy * 11) % 256, (xx * yy) % 256], -1)
    return np.clip(rgb + rng.integers(-25, 25, size=rgb.shape), 0, 255).astype(np.uint8)


@pytest.mark.parametrize("shape", [(1, 1), (8, 8), (16, 16), (17, 23), (33, 50), (40, 24), (47, 81)])
@pytest.mark.parametrize("quality", [30, 50, 75, 90, 100])
def test_byte_identical_to_libjpeg_turbo(shape, quality):
    rgb = _frame(*shape, seed=shape[0] * 100 + shape[1])
    assert J.encode(rgb, quality) == _pil(rgb, quality)


def test_dct_of_constant_block():
    for v in (-128, -1, 0, 5, 127):
        c = J.fdct_islow(np.full((8, 8), v))
        assert c[0, 0] == 64 * v and not c.reshape(-1)[1:].any()


def test_low_quality_uses_16_bit_tables():
    jpg = J.encode(_frame(16, 16, 1), 10)
    assert b"\xff\xdb\x00\x83\x10" in jpg and b"\xff\xc1" in jpg


def test_frame_message():
    rgba = np.zeros((8, 8, 4), np.uint8)
    msg = J.frame_message(rgba, 90)
    assert msg.startswith(b'{"payload": "data:image/jpeg;base64,/9j/4AAQSkZJRgABAQ') and msg.endswith(b'"}')
