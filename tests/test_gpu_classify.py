"""Classification (the combined RGBA8 textures) bit-exact against the oracle over
the source layouts that select each classification kernel: one or two active
sources with 1 or 3 components in either order (the kernels compiled per source
signature, four voxels per thread), and the layouts the generic kernel keeps:
rows that are not a multiple of 4 voxels, three active sources, a non-persistent
source behind a guard, an inactive source between active ones, every dither mode."""
import dataclasses

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu


def _sources(shape, seed):
    rng = np.random.default_rng(seed)
    Z, Y, X = shape
    dens = rng.uniform(0, 1, size=(Z, Y, X)).astype(np.float32)
    vec = rng.normal(size=(Z, Y, X, 3)).astype(np.float32)
    d = scenes.VolumeSourceSpec("density", dens, "idem", list(scenes.DENSITY_TF), 2.0, 0.5)
    v = scenes.VolumeSourceSpec("vector", vec, "length", list(scenes.VECTOR_TF), 1.5, 0.0)
    return d, v


def _spec(shape, vols, dither=1, name="classify"):
    Z, Y, X = shape
    return scenes.SceneSpec(name=name, global_size=(X, Y, Z), local_size=(X, Y, Z), particle_grid=(4, 4, 4),
                            volumes=list(vols), framebuffer=(32, 24), dither_mode=dither)


def _check(spec, gpu, guard=0):
    import torch

    from isaac_amd import visualization as V
    from oracle import oracle as O

    vis = V.from_scene(spec, gpu.index or 0)
    if guard:
        # the same sources behind a guard of `guard` voxels (rows no longer 16-byte aligned)
        dev = torch.device("cuda", gpu.index or 0)
        g = guard
        srcs = []
        for v in spec.volumes:
            pad = ((g, g), (g, g), (g, g)) + (((0, 0),) if v.data.ndim == 4 else ())
            t = torch.from_numpy(np.pad(v.data, pad, constant_values=-7.0)).to(dev)
            srcs.append(V.VolumeSource(v.name, t, v.persistent, guard_size=g))
        torch.cuda.synchronize(dev)
        vis.close()
        vis = V.IsaacVisualization(spec.name, spec.framebuffer, spec.global_size, spec.local_size, spec.particle_grid,
                                   spec.position, srcs, [], [], spec.scale, spec.transfer_size, gpu.index or 0)
        V.apply_scene(vis, spec)
    vis.update_buffers()
    vol, iso = vis.combined_textures()
    ora = O.render(spec, 32, want_textures=True)
    nv = int((vol != ora.vol_tex).any(axis=-1).sum())
    ni = int((iso != ora.iso_tex).any(axis=-1).sum())
    print(f"{spec.name}: {nv} volume / {ni} iso texels differ of {vol.shape[0] * vol.shape[1] * vol.shape[2]}", flush=True)
    assert nv == 0 and ni == 0
    assert vol[..., 3].any()
    vis.close()


LAYOUTS = {
    "density": lambda d, v: [d],
    "vector": lambda d, v: [v],
    "density+vector": lambda d, v: [d, v],
    "vector+density": lambda d, v: [v, d],
    "vector+vector": lambda d, v: [v, dataclasses.replace(v, name="vector2", chain="mul(0.5,2,1)|length", iso=0.3)],
    "three sources": lambda d, v: [d, v, dataclasses.replace(d, name="density2", chain="mul(0.8)", weight=0.5)],
    "inactive between": lambda d, v: [d, dataclasses.replace(d, name="off", weight=0.0, iso=0.0), v],
    "iso only": lambda d, v: [dataclasses.replace(d, weight=0.0, iso=0.4), v],
}


@pytest.mark.parametrize("layout", sorted(LAYOUTS))
@pytest.mark.parametrize("shape", [(24, 20, 32), (22, 18, 30)], ids=["x32", "x30"])
def test_layouts(gpu, oracle_lib, layout, shape):
    d, v = _sources(shape, 3)
    _check(_spec(shape, LAYOUTS[layout](d, v), name=f"{layout} {shape}"), gpu)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_dither_modes_four_voxels_per_thread(gpu, oracle_lib, mode):
    shape = (20, 18, 36)
    d, v = _sources(shape, 5)
    _check(_spec(shape, [d, v], dither=mode, name=f"dither {mode}"), gpu)


def test_non_persistent(gpu, oracle_lib):
    shape = (16, 16, 16)
    d, v = _sources(shape, 7)
    _check(_spec(shape, [dataclasses.replace(d, persistent=False), v], name="non-persistent"), gpu)


@pytest.mark.parametrize("persistent", [True, False])
def test_guarded_sources(gpu, oracle_lib, persistent):
    shape = (16, 12, 20)
    d, v = _sources(shape, 9)
    d = dataclasses.replace(d, persistent=persistent)
    _check(_spec(shape, [d, v], name=f"guard 1 persistent={persistent}"), gpu, guard=1)
