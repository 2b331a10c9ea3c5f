"""Field sources with advection on the GPU against the oracle: the noise texture
(Halton seeds, three 7-tap blur passes, clamp) and every frame's advection
texture bit-exact over several buffer updates (history, seeding window), the
combined textures bit-exact with the field source's alpha scaled by its
advection texture, and the frame within the pixel bar."""
import dataclasses

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu


def _spec(n=40, scale=(1.0, 1.0, 1.0), advection=(1.0, 0.95, 3, 1, 500)):
    base = scenes.prod_scene(n=n, fb=(160, 120), n_particles=2000)
    field = scenes.VolumeSourceSpec("flow", scenes.curl_noise((n, n, n), 7, 4), "length",
                                    [(0, (0.1, 0.4, 1.0, 0.0)), (300, (0.2, 0.8, 1.0, 0.6)), (1024, (1.0, 1.0, 1.0, 1.0))],
                                    2.0, 0.0)
    return dataclasses.replace(base, volumes=base.volumes[:1], fields=[field], scale=scale, advection=advection)


@pytest.mark.parametrize("scale", [(1.0, 1.0, 1.0), (1.0, 1.5, 0.75)])
def test_advection_sequence_bit_exact(gpu, oracle_lib, scale):
    from isaac_amd import visualization as V
    from oracle import oracle as O

    spec = _spec(scale=scale)
    step, weight, period, duration, seeds = spec.advection
    vis = V.from_scene(spec, gpu.index or 0)
    X, Y, Z = spec.local_size
    field = np.ascontiguousarray(spec.fields[0].data, np.float32)
    noise = O.noise(spec.local_size, scale, seeds)
    hist = np.zeros((Z, Y, X), np.uint8)
    for k in range(1, 6):
        vis.update_buffers()
        adv, noise_gpu, ts = vis.advection(0)
        assert ts == k
        if k == 1:
            assert np.array_equal(noise_gpu.view(np.uint32), noise.view(np.uint32)), np.abs(noise_gpu - noise).max()
        want = O.advect(field, noise, hist, scale, float(np.float32(step) * np.float32(3)), weight, (k % period) < duration)
        nd = int((adv != want).sum())
        print(f"scale {scale} step {k}: {nd} differing voxels, mean {adv.mean():.2f}", flush=True)
        assert nd == 0
        hist = want
    assert hist.any()
    # the classification uses the advection texture of the last update
    ps = O.PreparedScene(spec)
    O.set_advection(ps, {1: hist})
    ora = O.render(spec, 32, want_textures=True, prepared=ps)
    vol, iso = vis.combined_textures()
    assert np.array_equal(vol, ora.vol_tex)
    assert np.array_equal(iso, ora.iso_tex)
    img = vis.draw()
    d = np.abs(img.astype(np.int32) - ora.rgba.astype(np.int32)).max(axis=2)
    print(f"frame: {(d <= 2).mean() * 100:.4f}% within 2/255, max {d.max()}", flush=True)
    assert (d <= 2).mean() >= 0.999
    vis.close()


def test_seed_points_rebuild_noise(gpu, oracle_lib):
    from isaac_amd import visualization as V
    from oracle import oracle as O

    spec = _spec(n=24)
    vis = V.from_scene(spec, gpu.index or 0)
    vis.update_buffers()
    vis.seed_points = 77
    vis.update_buffers()
    _, noise_gpu, ts = vis.advection(0)
    assert ts == 2
    want = O.noise(spec.local_size, spec.scale, 77)
    assert np.array_equal(noise_gpu.view(np.uint32), want.view(np.uint32))
    vis.close()


def test_field_sources_rejected_by_brick_form(gpu):
    from isaac_amd import visualization as V
    from isaac_amd._lib import IsaacHipError

    spec = _spec(n=24)
    vis = V.from_scene(spec, gpu.index or 0)
    vis.set_render_form(vis.FORM_BRICKS, vis.F16)
    with pytest.raises(IsaacHipError):
        vis.update_buffers()
    vis.close()
