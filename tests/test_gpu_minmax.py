"""The min/max request on the GPU (isaac_hip_min_max) equals the oracle's
sequential restatement bit for bit: production-shaped sources (density, float3
field through `length`, particles), special values (NaN, +-inf, signed-zero
ties), non-persistent sources read from their texture, an empty particle source."""
import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.asarray(a, np.float32).view(np.uint32)


def _check(spec, gpu, update=True):
    from isaac_amd import visualization as V
    from oracle import oracle as O

    vis = V.from_scene(spec, gpu.index or 0)
    if update:
        vis.update_buffers()
    got = vis.min_max()
    want = O.min_max(spec)
    print(f"{spec.name}: min {got[0]} max {got[1]}", flush=True)
    assert np.array_equal(_bits(got[0]), _bits(want[0])), (got[0], want[0])
    assert np.array_equal(_bits(got[1]), _bits(want[1])), (got[1], want[1])
    vis.close()
    return got


def test_production_shaped_sources(gpu):
    spec = scenes.prod_scene(n=48, fb=(64, 64), n_particles=50_000)
    _check(spec, gpu)


def test_special_values(gpu):
    Z, Y, X = 24, 20, 16
    rng = np.random.default_rng(2)
    a = rng.uniform(-1, 1, size=(Z, Y, X)).astype(np.float32)
    a[rng.random(a.shape) < 0.05] = np.nan
    a[3, 4, 5] = np.inf
    a[7, 1, 2] = -np.inf
    b = np.abs(rng.uniform(0.5, 1, size=(Z, Y, X))).astype(np.float32)
    b[5, 3, 0] = -0.0  # the first zero in column order decides the sign of the min
    b[9, 3, 0] = 0.0
    b[0, 11, 7] = 0.0
    c = np.full((Z, Y, X), np.nan, np.float32)
    vols = [scenes.VolumeSourceSpec("a", a, "mul(-2)"), scenes.VolumeSourceSpec("b", b),
            scenes.VolumeSourceSpec("c", c)]
    spec = scenes.SceneSpec(name="special", global_size=(X, Y, Z), local_size=(X, Y, Z), particle_grid=(4, 4, 4),
                            volumes=vols, framebuffer=(32, 32))
    mins, maxs = _check(spec, gpu)
    assert np.signbit(mins[1]) and mins[1] == 0.0
    assert mins[2] == np.finfo(np.float32).max and maxs[2] == -np.finfo(np.float32).max


def test_non_persistent_sources_and_empty_particles(gpu):
    rng = np.random.default_rng(4)
    n = 32
    vec = rng.normal(size=(n, n, n, 3)).astype(np.float32)
    dens = rng.normal(size=(n, n, n)).astype(np.float32)
    vols = [scenes.VolumeSourceSpec("v", vec, "length", persistent=False),
            scenes.VolumeSourceSpec("d", dens, "add(0.25)|mul(3)", persistent=False)]
    parts = [scenes.ParticleSourceSpec("p", np.zeros((0, 3), np.float32), np.zeros((0, 1), np.float32))]
    spec = scenes.SceneSpec(name="nonpersistent", global_size=(n, n, n), local_size=(n, n, n),
                            particle_grid=(8, 8, 8), volumes=vols, particles=parts, framebuffer=(32, 32))
    _check(spec, gpu)


def test_large_volume_many_workgroups(gpu):
    """256^3 fields and 2 M particles: many voxels per thread of the reduction grid."""
    spec = scenes.prod_scene(n=256, fb=(64, 64), n_particles=2_000_000)
    _check(spec, gpu)
