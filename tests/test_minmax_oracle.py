"""The oracle's min/max request (oracle_min_max) against an independent numpy
statement of what the reference's loops return (MinMaxKernel and the host
reduction, lib/isaac/isaac_min_max_kernel.hpp:22-97, lib/isaac.hpp:372-497):

  * the minimum over the chain values that compare below FLT_MAX (NaN and +inf
    never replace the start value), the maximum over those above -FLT_MAX;
  * among equal values (+0.0 / -0.0) the first in visiting order: columns
    x + y*X in order, z inside a column, a cell's particles by index.
"""
import numpy as np
import pytest

from isaac_amd import scenes

FMAX = np.float32(np.finfo(np.float32).max)


def _expected(values_in_order: np.ndarray):
    """values_in_order: float32 chain values in the reference's visiting order."""
    v = np.asarray(values_in_order, np.float32)
    lo_ok = v < FMAX
    hi_ok = v > -FMAX
    if lo_ok.any():
        m = v[lo_ok].min()
        mn = v[lo_ok][np.flatnonzero(v[lo_ok] == m)[0]]
    else:
        mn = FMAX
    if hi_ok.any():
        m = v[hi_ok].max()
        mx = v[hi_ok][np.flatnonzero(v[hi_ok] == m)[0]]
    else:
        mx = -FMAX
    return np.float32(mn), np.float32(mx)


def _column_order(a: np.ndarray) -> np.ndarray:
    """(Z, Y, X) -> values in the order columns (x + y*X), then z."""
    Z, Y, X = a.shape
    return np.transpose(a, (1, 2, 0)).reshape(Y * X * Z)


def _spec(vols, parts=(), local=(6, 5, 4), grid=(2, 2, 2)):
    return scenes.SceneSpec(name="mm", global_size=local, local_size=local, particle_grid=grid,
                            volumes=list(vols), particles=list(parts), framebuffer=(8, 8))


def _same_bits(a, b):
    return np.asarray(a, np.float32).view(np.uint32).tolist() == np.asarray(b, np.float32).view(np.uint32).tolist()


def test_scalar_and_vector_chains():
    from oracle import oracle as O

    rng = np.random.default_rng(3)
    Z, Y, X = 4, 5, 6
    dens = rng.normal(size=(Z, Y, X)).astype(np.float32)
    vec = rng.normal(size=(Z, Y, X, 3)).astype(np.float32)
    spec = _spec([scenes.VolumeSourceSpec("d", dens, "mul(2)|add(1)"),
                  scenes.VolumeSourceSpec("v", vec, "length")])
    mins, maxs = O.min_max(spec)
    e0 = _expected(_column_order((dens * np.float32(2)) + np.float32(1)))
    n2 = (vec[..., 0] * vec[..., 0] + vec[..., 1] * vec[..., 1]) + vec[..., 2] * vec[..., 2]
    e1 = _expected(_column_order(np.sqrt(n2.astype(np.float32))))
    assert _same_bits(mins, [e0[0], e1[0]]) and _same_bits(maxs, [e0[1], e1[1]])


def test_special_values_and_signed_zero_ties():
    from oracle import oracle as O

    Z, Y, X = 4, 5, 6
    a = np.full((Z, Y, X), 0.5, np.float32)
    a[1, 2, 3] = np.nan
    a[0, 0, 0] = np.inf  # never the min (inf < FLT_MAX fails), the max
    a[3, 4, 5] = -np.inf  # the min, never the max
    spec = _spec([scenes.VolumeSourceSpec("s", a)])
    mins, maxs = O.min_max(spec)
    assert mins[0] == -np.inf and maxs[0] == np.inf
    # ties of +0 and -0: the first in column order decides the sign
    b = np.full((Z, Y, X), 1.0, np.float32)
    b[2, 0, 0] = -0.0  # column 0, z = 2
    b[0, 1, 0] = 0.0  # column X (later)
    b[3, 0, 0] = 0.0  # column 0, z = 3 (later)
    mins, _ = O.min_max(_spec([scenes.VolumeSourceSpec("s", b)]))
    assert mins[0] == 0.0 and np.signbit(mins[0])
    b[1, 0, 0] = 0.0  # column 0, z = 1: now +0 comes first
    mins, _ = O.min_max(_spec([scenes.VolumeSourceSpec("s", b)]))
    assert mins[0] == 0.0 and not np.signbit(mins[0])
    # all NaN: the start values survive
    c = np.full((Z, Y, X), np.nan, np.float32)
    mins, maxs = O.min_max(_spec([scenes.VolumeSourceSpec("s", c)]))
    assert mins[0] == FMAX and maxs[0] == -FMAX


def test_particles_visit_cells_in_column_order():
    from oracle import oracle as O

    rng = np.random.default_rng(5)
    local, grid = (8, 8, 8), (2, 2, 2)
    pos = rng.uniform(0, 8, size=(200, 3)).astype(np.float32)
    attr = rng.normal(size=(200, 2)).astype(np.float32)
    attr[17, :] = 0.0
    attr[150, :] = [-0.0, 0.0]
    part = scenes.ParticleSourceSpec("p", pos, attr, chain="idem")
    vol = scenes.VolumeSourceSpec("d", np.zeros((8, 8, 8), np.float32), weight=0.0)
    mins, maxs = O.min_max(_spec([vol], [part], local, grid))
    cw = np.float32(8) / np.float32(2)
    cells = np.clip(np.floor(pos / cw), 0, 1).astype(np.int64)
    key = ((cells[:, 0] + cells[:, 1] * 2) * 2 + cells[:, 2]) * len(pos) + np.arange(len(pos))
    order = np.argsort(key, kind="stable")
    e = _expected(attr[order, 0])
    assert _same_bits(mins[1:], [e[0]]) and _same_bits(maxs[1:], [e[1]])


def test_empty_particle_source():
    from oracle import oracle as O

    part = scenes.ParticleSourceSpec("p", np.zeros((0, 3), np.float32), np.zeros((0, 1), np.float32))
    vol = scenes.VolumeSourceSpec("d", np.ones((4, 5, 6), np.float32))
    mins, maxs = O.min_max(_spec([vol], [part]))
    assert mins[1] == FMAX and maxs[1] == -FMAX
    assert mins[0] == 1.0 and maxs[0] == 1.0


@pytest.mark.parametrize("chain", ["pow(2)", "sum", "add(-3)|mul(0.5)"])
def test_more_chains(chain):
    from oracle import oracle as O

    rng = np.random.default_rng(11)
    vec = rng.uniform(-2, 2, size=(4, 5, 6, 2)).astype(np.float32)
    mins, maxs = O.min_max(_spec([scenes.VolumeSourceSpec("v", vec, chain)]))
    if chain == "pow(2)":
        vals = np.power(vec[..., 0], np.float32(2))
    elif chain == "sum":
        vals = vec[..., 0] + vec[..., 1]
    else:
        vals = (vec[..., 0] + np.float32(-3)) * np.float32(0.5)
    e = _expected(_column_order(vals.astype(np.float32)))
    assert _same_bits(mins, [e[0]]) and _same_bits(maxs, [e[1]])
