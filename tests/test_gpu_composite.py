"""Sort-last compositing on the GPU (isaac_hip_composite): bit-exact against the
oracle's pairwise tree of IceT byte "over" for 1..64 images, tails of 1..3
pixels and full frames; and, end to end, a volume split into 2x2x2 sub-volumes
rendered by eight contexts with one camera, ordered by isaac_hip_composite_order
and composited, against the same volume rendered whole (differences: the
sub-volume borders, where trilinear taps clamp without guard cells, and the
RGBA8 rounding of each sub-volume's partial colour)."""
import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu


def _premul(rng, n):
    a = rng.integers(0, 256, size=(n, 1))
    rgb = (rng.random((n, 3)) * (a + 1)).astype(np.int64).clip(0, a)
    return np.concatenate([rgb, a], axis=1).astype(np.uint8)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 33, 64])
@pytest.mark.parametrize("pixels", [1, 7, 4099, 1920 * 1080])
def test_kernel_bit_exact(gpu, n, pixels):
    import torch

    from isaac_amd import composite as K
    from oracle import composite_oracle as C

    if pixels == 1920 * 1080 and n not in (8, 64):
        pytest.skip("full frames for 8 and 64 images only")
    rng = np.random.default_rng(n * 7919 + pixels)
    imgs = [_premul(rng, pixels) for _ in range(n)]
    order = [int(v) for v in rng.permutation(n)]
    dev = [torch.from_numpy(im).to(gpu) for im in imgs]
    out = K.composite(dev, order)
    torch.cuda.synchronize()
    got = out.cpu().numpy()
    want = C.composite(imgs, order)
    nd = int((got != want).any(axis=-1).sum())
    print(f"n={n} pixels={pixels}: {nd} pixels differ", flush=True)
    assert nd == 0


def test_rejects_bad_arguments(gpu):
    import torch

    from isaac_amd import composite as K
    from isaac_amd._lib import IsaacHipError

    a = torch.zeros((16, 4), dtype=torch.uint8, device=gpu)
    b = torch.zeros((16, 4), dtype=torch.uint8, device=gpu)
    with pytest.raises(IsaacHipError):
        K.composite([a, b], [0, 0])  # not a permutation
    with pytest.raises(IsaacHipError):
        K.composite([a, b], [0, 2])
    with pytest.raises(IsaacHipError):
        K.composite([a[1:], b[1:]], [0, 1])  # not 16-byte aligned


def _psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)


def test_split_volume_composites_to_the_whole(gpu):
    import torch

    from isaac_amd import composite as K
    from isaac_amd import visualization as V

    n, h = 64, 32
    g = scenes.gaussian_mixture((n, n, n), 2, 8, (4.0, 12.0))
    dens = np.clip(g / max(float(g.max()), 1e-6), 0, 1).astype(np.float32)
    a, b = 0.5, 0.6
    ry = np.array([[np.cos(a), 0.0, np.sin(a)], [0.0, 1.0, 0.0], [-np.sin(a), 0.0, np.cos(a)]])
    rx = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(b), -np.sin(b)], [0.0, np.sin(b), np.cos(b)]])
    cam = scenes.Camera(distance=-4.0, rotation=rx @ ry)
    base = dict(global_size=(n, n, n), particle_grid=(4, 4, 4), framebuffer=(256, 256), step=0.5,
                interpolation=True, camera=cam, background=(0.0, 0.0, 0.0, 0.0))

    def spec_for(pos, size, data):
        # an iso-surface (opaque) over the volume: the composite order decides what is in front
        vol = scenes.VolumeSourceSpec("density", data, "idem", list(scenes.DENSITY_TF), 2.0, 0.5)
        return scenes.SceneSpec(name=f"part{pos}", local_size=size, position=pos, volumes=[vol], **base)

    whole = V.from_scene(spec_for((0, 0, 0), (n, n, n), dens), gpu.index or 0)
    whole.update_buffers()
    ref = whole.draw()
    mv = whole.modelview
    whole.close()
    positions, sizes, images = [], [], []
    for z in range(2):
        for y in range(2):
            for x in range(2):
                pos = (x * h, y * h, z * h)
                part = np.ascontiguousarray(dens[z * h:(z + 1) * h, y * h:(y + 1) * h, x * h:(x + 1) * h])
                vis = V.from_scene(spec_for(pos, (h, h, h), part), gpu.index or 0)
                assert np.allclose(vis.modelview, mv)
                vis.update_buffers()
                images.append(torch.from_numpy(vis.draw()).to(gpu))
                vis.close()
                positions.append(pos)
                sizes.append((h, h, h))
    order = K.composite_order(positions, sizes, (n, n, n), (1.0, 1.0, 1.0), mv)
    out = K.composite(images, order).cpu().numpy()
    p = _psnr(out, ref)
    shuffled = K.composite(images, list(reversed(order))).cpu().numpy()
    ps = _psnr(shuffled, ref)
    print(f"split 2x2x2: {p:.2f} dB against the whole render (reversed order: {ps:.2f} dB); order {order}", flush=True)
    assert (ref[..., 3] > 0).mean() > 0.1
    assert p >= 30.0 and p > ps + 6.0
