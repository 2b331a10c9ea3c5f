"""The oracle's field-source advection pinned against independent statements:

  * halton: the radical inverse i -> sum of digits(i, base) * base^-(k+1);
  * the noise texture: a numpy statement of three separable 7-tap Gauss passes
    as shifted sums (np.roll for REPEAT, the sampler's half-voxel offset as the
    average of two neighbours), then clamp(30 v, 0, 1);
  * the advection step: a zero field keeps the history (only the sampler centre
    tap contributes), a unit +x field with a one-voxel step shifts it by one
    voxel with zeros entering at the border, the seeded noise is dithered with
    the 16^3 mask.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _halton_exact(i, b):
    r, f = Fraction(0), Fraction(1)
    while i > 0:
        f /= b
        r += f * (i % b)
        i //= b
    return r


def test_noise_seed_positions_follow_the_halton_sequence(oracle_lib):
    from oracle import oracle as O

    X, Y, Z = 19, 23, 29
    # few seed points: each lands where float(X) * halton(i, 3) etc. truncate
    for count in (1, 7, 40):
        n = O.noise((X, Y, Z), (1.0, 1.0, 1.0), count)
        seeds = np.zeros((Z, Y, X), np.float64)
        for i in range(1, count + 1):
            x = int(X * float(_halton_exact(i, 3)))
            y = int(Y * float(_halton_exact(i, 5)))
            z = int(Z * float(_halton_exact(i, 7)))
            seeds[z, y, x] = 1.0
        want = _noise_numpy(seeds, (1.0, 1.0, 1.0))
        assert np.abs(n - want).max() <= 2e-6, np.abs(n - want).max()


def _blur_axis(a, axis, scale):
    """Sampler LINEAR/REPEAT at coord + k/scale along `axis` (scale 1, 2 or 0.5 here),
    weights 64, 40.5, 10, 1 (|k| = 0..3), / 167.  The sampler is 3-D and takes
    coord - 0.5: along the two other axes every pass averages voxels c-1 and c."""
    w = {0: 64.0, 1: 40.5, 2: 10.0, 3: 1.0}
    for other in (0, 1, 2):
        if other != axis:
            a = 0.5 * (a + np.roll(a, 1, axis=other))
    out = np.zeros_like(a)
    for k in range(-3, 4):
        off = k / scale - 0.5  # the sampler's half-voxel shift
        lo = int(np.floor(off))
        t = off - lo
        s = (1 - t) * np.roll(a, -lo, axis=axis) + t * np.roll(a, -(lo + 1), axis=axis)
        out += w[abs(k)] * s
    return out / 167.0


def _noise_numpy(seeds, scale):
    a = _blur_axis(seeds, 2, scale[0])
    a = _blur_axis(a, 1, scale[1])
    a = _blur_axis(a, 0, scale[2])
    return np.clip(a * 30.0, 0.0, 1.0)


@pytest.mark.parametrize("scale", [(1.0, 1.0, 1.0), (2.0, 1.0, 0.5)])
def test_noise_blur_matches_shifted_sums(oracle_lib, scale):
    from oracle import oracle as O

    X, Y, Z = 24, 20, 16
    count = 300
    n = O.noise((X, Y, Z), scale, count)
    seeds = np.zeros((Z, Y, X))
    for i in range(1, count + 1):
        seeds[int(Z * float(_halton_exact(i, 7))), int(Y * float(_halton_exact(i, 5))),
              int(X * float(_halton_exact(i, 3)))] = 1.0
    want = _noise_numpy(seeds, scale)
    assert n.dtype == np.float32 and n.shape == (Z, Y, X)
    assert np.abs(n - want).max() <= 2e-6
    assert n.max() > 0.3 and n.min() >= 0.0


def _mask16():
    t = json.load(open(os.path.join(GOLDEN, "dither_tables.json")))["tables"]["4"]
    return np.array(t["mask"], np.float64).reshape(16, 16, 16) / 4096.0  # [z][y][x]


def _dither16(v):
    """(Z, Y, X) float32 values -> bytes with the 16^3 blue-noise mask."""
    Z, Y, X = v.shape
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    d = _mask16()[z % 16, y % 16, x % 16].astype(np.float32)
    s = v.astype(np.float32) * np.float32(255)
    i = s.astype(np.int64) + (s - np.floor(s) > d)
    return np.clip(i, 0, 255).astype(np.uint8)


def test_zero_field_keeps_history(oracle_lib):
    from oracle import oracle as O

    rng = np.random.default_rng(1)
    Z, Y, X = 10, 12, 14
    hist = rng.integers(0, 256, size=(Z, Y, X)).astype(np.uint8)
    field = np.zeros((Z, Y, X, 3), np.float32)
    out = O.advect(field, np.zeros((Z, Y, X), np.float32), hist, (1, 1, 1), 3.0, 1.0, False)
    assert np.array_equal(out, hist)


def test_unit_field_shifts_history_by_one_voxel(oracle_lib):
    from oracle import oracle as O

    rng = np.random.default_rng(2)
    Z, Y, X = 10, 12, 14
    hist = rng.integers(0, 256, size=(Z, Y, X)).astype(np.uint8)
    field = np.zeros((Z, Y, X, 3), np.float32)
    field[..., 0] = 2.5  # any positive length: the unit direction is +x
    step = np.float32(1.0 / 3.0) * np.float32(3.0)  # step factor 1/3 -> one voxel
    out = O.advect(field, np.zeros((Z, Y, X), np.float32), hist, (1, 1, 1), float(step), 1.0, False)
    want = np.zeros_like(hist)
    want[:, :, 1:] = hist[:, :, :-1]
    assert np.array_equal(out, want)
    # the weight scales the history before dithering
    out2 = O.advect(field, np.zeros((Z, Y, X), np.float32), hist, (1, 1, 1), float(step), 0.5, False)
    v = want.astype(np.float32) / np.float32(255) * np.float32(0.5)
    assert np.array_equal(out2, _dither16(v))


def test_seeding_dithers_the_noise(oracle_lib):
    from oracle import oracle as O

    rng = np.random.default_rng(3)
    Z, Y, X = 20, 18, 17
    noise = rng.uniform(0, 1, size=(Z, Y, X)).astype(np.float32)
    field = rng.normal(size=(Z, Y, X, 3)).astype(np.float32)
    zero = np.zeros((Z, Y, X), np.uint8)
    out = O.advect(field, noise, zero, (1, 1, 1), 3.0, 0.95, True)
    assert np.array_equal(out, _dither16(noise))
    out = O.advect(field, noise, zero, (1, 1, 1), 3.0, 0.95, False)
    assert not out.any()
