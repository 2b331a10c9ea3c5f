"""Sort-last compositing (one GPU's share of the reference's IceT step): the
oracle's composite order and byte blending pinned by what the algebra fixes, and
the library's host-side composite order equal to the oracle's.

  * order: sub-volumes on the view axis come nearest first; mirror-symmetric
    sub-volumes tie and keep index order (std::multimap); a 2x2x2 split seen from
    a corner puts the corner's sub-volume first and the opposite one last;
  * over: an opaque front replaces the back, an empty front (all zero) keeps the
    back, an empty back keeps the front, alpha 0/255 operands are exact, results
    of premultiplied operands stay premultiplied;
  * the pairwise tree equals a level-by-level binary-swap simulation for
    power-of-two counts and is the sequential blend whenever every alpha is 0 or
    255 (then "over" is exact and associative).
"""
import numpy as np
import pytest

from oracle import composite_oracle as C


def _mv(distance, rot=np.eye(3)):
    m = np.eye(4)
    m[:3, :3] = rot
    m[2, 3] = distance
    return m.T.reshape(16)  # column-major


def _split(n_side, size):
    h = size // n_side
    pos, loc = [], []
    for z in range(n_side):
        for y in range(n_side):
            for x in range(n_side):
                pos.append((x * h, y * h, z * h))
                loc.append((h, h, h))
    return pos, loc


def test_order_along_view_axis():
    # sub-volumes stacked in z; the camera looks down -z from +z (distance -4.5 after the modelview)
    pos = [(0, 0, 64 * k) for k in range(4)]
    loc = [(64, 64, 64)] * 4
    order = C.composite_order(pos, loc, (64, 64, 256), (1, 1, 1), _mv(-4.5))
    assert order == [3, 2, 1, 0]
    order = C.composite_order(pos, loc, (64, 64, 256), (1, 1, 1), _mv(-4.5, np.diag([1.0, -1.0, -1.0])))
    assert order == [0, 1, 2, 3]


def test_order_ties_keep_index_order():
    pos, loc = _split(2, 64)
    # looking straight down z: the four sub-volumes of a z layer are at equal distance
    order = C.composite_order(pos, loc, (64, 64, 64), (1, 1, 1), _mv(-4.5))
    assert order == [4, 5, 6, 7, 0, 1, 2, 3]


def test_order_corner_view():
    pos, loc = _split(2, 64)
    c, s = np.cos(np.pi / 4), np.sin(np.pi / 4)
    ry = np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    rx = np.array([[1, 0, 0], [0, c, -s], [0, s, c]])
    order = C.composite_order(pos, loc, (64, 64, 64), (1, 1, 1), _mv(-4.5, rx @ ry))
    d = [np.linalg.norm((rx @ ry) @ ((np.array(p) + 16.0 - 32.0) / 32.0) + np.array([0, 0, -4.5])) for p in pos]
    assert order[0] == int(np.argmin(d)) and order[-1] == int(np.argmax(d))


def test_library_order_equals_oracle(hip_lib):
    from isaac_amd import composite as K

    rng = np.random.default_rng(3)
    for trial in range(20):
        n_side = int(rng.integers(1, 4))
        pos, loc = _split(n_side, 96)
        a, b = rng.uniform(0, 2 * np.pi, 2)
        ry = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
        rx = np.array([[1, 0, 0], [0, np.cos(b), -np.sin(b)], [0, np.sin(b), np.cos(b)]])
        scale = tuple(rng.choice([0.5, 1.0, 1.5], 3))
        mv = _mv(-float(rng.uniform(2, 6)), rx @ ry)
        assert K.composite_order(pos, loc, (96, 96, 96), scale, mv) == C.composite_order(pos, loc, (96, 96, 96), scale, mv)


def _premul(rng, shape):
    a = rng.integers(0, 256, size=shape[:-1] + (1,))
    rgb = (rng.random(shape[:-1] + (3,)) * (a + 1)).astype(np.int64).clip(0, a)
    return np.concatenate([rgb, a], axis=-1).astype(np.uint8)


def test_over_identities():
    rng = np.random.default_rng(1)
    back = _premul(rng, (1000, 4))
    front = _premul(rng, (1000, 4))
    opaque = front.copy()
    opaque[:, 3] = 255
    assert np.array_equal(C.over_ubyte(opaque, back), opaque)
    assert np.array_equal(C.over_ubyte(np.zeros_like(front), back), back)
    assert np.array_equal(C.over_ubyte(front, np.zeros_like(back)), front)
    out = C.over_ubyte(front, back)
    assert (out[:, :3] <= out[:, 3:4]).all()  # premultiplied stays premultiplied
    # the textbook formula on bytes: back * (255 - a) / 255 rounded down, plus front
    want = (back.astype(np.int64) * (255 - front[:, 3:4].astype(np.int64))) // 255 + front
    assert np.array_equal(out, want.astype(np.uint8))


def _bswap(images):
    """binary swap for 2^k processes in composite order: in round r the partners
    i and i ^ 2^r exchange, the one earlier in the order is in front."""
    imgs = [np.asarray(i) for i in images]
    n = len(imgs)
    r = 1
    while r < n:
        imgs = [C.over_ubyte(imgs[i], imgs[i + r]) if (i // r) % 2 == 0 else imgs[i] for i in range(n)]
        r *= 2
    return imgs[0]


@pytest.mark.parametrize("n", [1, 2, 4, 8, 16])
def test_tree_is_binary_swap(n):
    rng = np.random.default_rng(n)
    imgs = [_premul(rng, (64, 4)) for _ in range(n)]
    order = list(rng.permutation(n))
    assert np.array_equal(C.composite(imgs, order), _bswap([imgs[o] for o in order]))


@pytest.mark.parametrize("n", [3, 5, 6, 7, 11])
def test_tree_is_sequential_when_exact(n):
    rng = np.random.default_rng(10 + n)
    imgs = []
    for _ in range(n):
        im = _premul(rng, (256, 4))
        a = np.where(rng.random((256, 1)) < 0.5, 0, 255)
        im = np.where(a == 0, 0, np.concatenate([im[:, :3], np.full((256, 1), 255)], axis=1)).astype(np.uint8)
        imgs.append(im)
    order = list(rng.permutation(n))
    seq = imgs[order[-1]]
    for o in reversed(order[:-1]):
        seq = C.over_ubyte(imgs[o], seq)
    assert np.array_equal(C.composite(imgs, order), seq)
