"""Pins for the CPU oracle (oracle/isaac_oracle.hpp) and the host-side restatements.

The reference ships no tests, golden images or fixtures, so every stage of the
oracle is pinned here against what the mathematics fixes: closed forms,
textbook algorithms evaluated independently in float64 numpy, brute force on
tiny inputs, and hand-derived tables from reading the reference (fixtures in
tests/golden/).  Each check is chosen so that a plausible transcription slip
(a dropped term, a wrong sign, a swapped axis, an off-by-one) fails it.
"""
import ctypes
import dataclasses
import json
import math
import os

import numpy as np
import pytest

from isaac_amd import scenes

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# --------------------------------------------------------------------------- transfer functions
def _closed_form_lut(points, T):
    """Piecewise-linear closed form in float64 over segments [p_k, p_{k+1})."""
    d = {}
    for k, c in points:
        d.setdefault(int(k), np.asarray(c, np.float64))
    keys = sorted(d)
    lut = np.zeros((T, 4))
    for a, b in zip(keys, keys[1:]):
        for j in range(a, min(b, T)):
            t = (j - a) / (b - a)
            lut[j] = d[a] * (1 - t) + d[b] * t
    return lut


@pytest.mark.parametrize(
    "points,T",
    [
        ([(0, (0, 0, 0, 0)), (1024, (1, 0.5, 0.25, 1))], 1024),
        (scenes.SMALL_TF, 1024),
        (scenes.DENSITY_TF, 1024),
        ([(10, (1, 1, 1, 1)), (20, (0, 0.5, 1, 0))], 32),  # entries outside [10,20) stay 0
        ([(0, (0, 0, 0, 0)), (64, (1, 1, 1, 1))], 32),  # a point beyond T: only j < T written
        ([(0, (1, 0, 0, 1)), (0, (0, 1, 0, 1)), (8, (0, 0, 1, 0))], 16),  # duplicate key: first wins
    ],
)
def test_transfer_lut_closed_form(oracle_lib, hip_lib, points, T):
    from isaac_amd import visualization as V

    ref = _closed_form_lut(points, T)
    ours = oracle_lib.build_lut(points, T)
    assert np.allclose(ours, ref, atol=2e-7, rtol=1e-6)
    prod = V.build_transfer_lut(points, T)
    assert np.array_equal(prod.view(np.uint32), ours.view(np.uint32)), "product LUT build differs from the restatement"
    # index T is never written; with the first point at p0 > 0 the head stays zero
    first = min(int(k) for k, _ in points)
    assert np.all(ours[:first] == 0)


def test_transfer_duplicate_key_first_wins(oracle_lib):
    lut = oracle_lib.build_lut([(0, (1, 0, 0, 1)), (0, (0, 1, 0, 1)), (8, (0, 0, 1, 0))], 16)
    assert np.allclose(lut[0], (1, 0, 0, 1))


def test_default_transfer_matches_reference_hsv(hip_lib):
    """Default TF of source i of n: hue 2*pi*i/n, alpha ramp 0->1 (volume) or 1 (particle)."""
    import colorsys

    n = 3
    for i in range(n):
        for part in (0, 1):
            pos = np.zeros(2, np.uint32)
            rgba = np.zeros(8, np.float32)
            hip_lib.isaac_hip_default_transfer(i, n, part, 1024, pos.ctypes.data, rgba.ctypes.data)
            r, g, b = colorsys.hsv_to_rgb(i / n, 1.0, 1.0)
            assert np.allclose(rgba[:3], (r, g, b), atol=1e-5)
            assert rgba[3] == (1.0 if part else 0.0) and rgba[7] == 1.0
            assert list(pos) == [0, 1024]


# --------------------------------------------------------------------------- functor chains
def _kat():
    return json.load(open(os.path.join(GOLDEN, "functor_chain_kat.json")))["cases"]


@pytest.mark.parametrize("case", _kat(), ids=lambda c: repr(c[0]))
def test_functor_parser_kat(oracle_lib, hip_lib, case):
    from isaac_amd import visualization as V

    chain, bc, params, err = case
    for parse in (oracle_lib.parse_chain, V.parse_functor_chain):
        got_bc, got_p, got_err = parse(chain)
        assert got_err == err, (parse, chain)
        assert list(got_bc) == bc, (parse, chain)
        assert np.allclose(np.asarray(got_p, np.float32), np.asarray(params, np.float32)), (parse, chain)


@pytest.mark.parametrize(
    "chain,value,expected",
    [
        ("idem", (1.0, 2.0, 3.0), 1.0),  # a vector chain ends in .x
        ("length", (3.0, 4.0, 0.0), 5.0),
        ("length", (1.0, 2.0, 2.0, 4.0), 5.0),
        ("mul(2)|add(1)", (3.0,), 7.0),
        ("mul(1,2,3)|sum", (1.0, 1.0, 1.0), 6.0),
        ("pow(2)", (3.0,), 9.0),
        ("sum|mul(3)", (1.0, 2.0), 9.0),
        ("add(1,2,3,4)|length", (0.0, 0.0, 0.0, 0.0), math.sqrt(30.0)),
        ("mul(-1)|add(2)|pow(3)", (1.0,), 1.0),
        ("add(0,5)|mul(1,2)|sum", (1.0, 1.0), 13.0),
    ],
)
def test_functor_chain_values(oracle_lib, chain, value, expected):
    bc, p, err = oracle_lib.parse_chain(chain)
    assert err == 0
    b = (ctypes.c_int32 * 3)(*bc)
    pr = (ctypes.c_float * 12)(*[x for row in p for x in row])
    v = (ctypes.c_float * 4)(*(list(value) + [0.0] * (4 - len(value))))
    got = oracle_lib.lib().oracle_apply_chain(b, pr, v, len(value))
    assert abs(got - expected) <= 1e-5 * max(1.0, abs(expected))


# --------------------------------------------------------------------------- dither
def _dither(oracle_lib, mode, x, y, z, c):
    out = (ctypes.c_uint8 * 4)()
    cc = (ctypes.c_float * 4)(*c)
    oracle_lib.lib().oracle_dither(mode, x, y, z, cc, out)
    return list(out)


def test_dither_mode0_rounds_half_away(oracle_lib):
    c = (np.float32(0.5) / np.float32(255), 1.2, -0.1, 0.25)
    assert _dither(oracle_lib, 0, 0, 0, 0, c) == [1, 255, 0, 64]


def _dither_tables():
    return json.load(open(os.path.join(GOLDEN, "dither_tables.json")))["tables"]


def test_dither_tables_are_intact():
    """The committed masks hash to what extraction recorded; the blue-noise masks
    (modes 2-4) are permutations of 0..n-1, so a mistyped entry shows up as a
    duplicate; the generated include files carry the same numbers."""
    import hashlib
    import re

    tables = _dither_tables()
    for mode, t in tables.items():
        vals = t["mask"]
        assert len(vals) == t["side"] ** 3
        assert hashlib.sha256(",".join(map(str, vals)).encode()).hexdigest() == t["sha256"], mode
        if mode != "1":
            assert sorted(vals) == list(range(len(vals))), mode
    assert sorted(tables["1"]["mask"])[0] == 1 and max(tables["1"]["mask"]) == 64
    root = os.path.dirname(os.path.dirname(GOLDEN))
    for inc in ("isaac_amd/csrc/dither_tables.inc", "oracle/dither_tables.inc"):
        txt = open(os.path.join(root, inc)).read()
        for mode, t in tables.items():
            body = re.search(r"DITHER_TABLE\(\w+, " + t["name"] + r", \d+\) = \{([^}]*)\}", txt).group(1)
            assert [int(v) for v in body.replace("\n", " ").split(",") if v.strip()] == t["mask"], (inc, mode)


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 7])
def test_dither_modes_against_fixture_tables(oracle_lib, mode):
    """Each texel rounds down just below its mask threshold and up just above it;
    the mask index is x % s + (y % s) * s + (z % s) * s^2 (s = 4, 4, 8, 16; modes
    other than 0-3 use the 16^3 mask)."""
    t = _dither_tables()[str(min(mode, 4))]
    side, div, mask = t["side"], t["divisor"], t["mask"]
    rng = np.random.default_rng(mode)
    for _ in range(300):
        x, y, z = (int(v) for v in rng.integers(0, 3 * side, size=3))
        m = mask[x % side + (y % side) * side + (z % side) * side * side]
        below = (10.0 + m / div - 0.25 / div) / 255.0
        got = _dither(oracle_lib, mode, x, y, z, (below, below, below, below))
        assert got == [10] * 4, (x, y, z, m, got)
        above = (10.0 + m / div + 0.25 / div) / 255.0
        got = _dither(oracle_lib, mode, x, y, z, (above, 0.0, 1.0, 2.0))
        assert got[0] == 11, (x, y, z, m, got)
        assert got[1:] == [0, 255, 255]


# --------------------------------------------------------------------------- sampler
def _tex():
    X = Y = Z = 8
    z, y, x = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    t = np.stack([10 + 3 * x + 5 * y + 7 * z, 200 - 2 * x - 3 * y - 4 * z, 5 * z, np.full_like(x, 100)], -1)
    return np.ascontiguousarray(t.astype(np.uint8)), (X, Y, Z)


def _sample(oracle_lib, tex, dims, p, linear):
    out = (ctypes.c_float * 4)()
    pp = (ctypes.c_float * 3)(*p)
    oracle_lib.lib().oracle_sample_b4(
        ctypes.c_void_p(tex.ctypes.data), ctypes.c_int64(dims[0]), ctypes.c_int64(dims[1]), ctypes.c_int64(dims[2]), pp, linear, out
    )
    return np.array(out)


def test_trilinear_exact_on_affine_field(oracle_lib):
    tex, dims = _tex()
    rng = np.random.default_rng(0)
    for _ in range(200):
        p = rng.uniform(0.5, 7.5, size=3)
        q = p - 0.5
        exp = np.array([10 + 3 * q[0] + 5 * q[1] + 7 * q[2], 200 - 2 * q[0] - 3 * q[1] - 4 * q[2], 5 * q[2], 100]) / 255
        assert np.allclose(_sample(oracle_lib, tex, dims, p, 1), exp, atol=2e-6)


def test_trilinear_clamps_at_borders(oracle_lib):
    tex, dims = _tex()
    # below 0.5 on x: both x taps clamp to 0 -> value of x = 0
    got = _sample(oracle_lib, tex, dims, (-3.0, 2.5, 3.5), 1)
    assert np.allclose(got * 255, [10 + 5 * 2 + 7 * 3, 200 - 3 * 2 - 4 * 3, 15, 100], atol=1e-4)
    got = _sample(oracle_lib, tex, dims, (20.0, 2.5, 3.5), 1)
    assert np.allclose(got * 255, [10 + 21 + 10 + 21, 200 - 14 - 6 - 12, 15, 100], atol=1e-4)


def test_nearest_truncates(oracle_lib):
    tex, dims = _tex()
    got = _sample(oracle_lib, tex, dims, (2.7, 3.2, -0.5), 0) * 255
    assert np.allclose(got, tex[0, 3, 2], atol=1e-4)
    got = _sample(oracle_lib, tex, dims, (-1.5, 7.99, 9.0), 0) * 255
    assert np.allclose(got, tex[7, 7, 0], atol=1e-4)


# --------------------------------------------------------------------------- camera
def test_perspective_closed_form(oracle_lib, hip_lib):
    from isaac_amd import visualization as V

    out = (ctypes.c_double * 16)()
    oracle_lib.lib().oracle_perspective(45.0, 1920 / 1080, float(np.float32(0.002)), 100.0, out)
    t = math.tan(22.5)  # glm 0.9.9 takes radians: the reference's 45.0 is 45 rad
    exp = np.zeros(16)
    exp[0], exp[5] = 1 / ((1920 / 1080) * t), 1 / t
    exp[10] = -(100 + 0.002) / (100 - 0.002)
    exp[11] = -1
    exp[14] = -(2 * 100 * 0.002) / (100 - 0.002)
    assert np.allclose(np.array(out), exp, rtol=1e-6)
    prod = V.default_projection(1920, 1080)
    assert np.array_equal(prod, np.array(out))
    assert np.array_equal(prod, oracle_lib.default_projection(1920, 1080))


def test_modelview_matches_restatement(oracle_lib, hip_lib):
    from isaac_amd import visualization as V

    R = scenes.rotation_axis((1, 2, 3), 37.0)
    a = V.modelview((0.1, -0.2, 0.3), R, -3.3)
    b = oracle_lib.modelview((0.1, -0.2, 0.3), R, -3.3)
    assert np.allclose(a, b, atol=1e-14)
    # distance translates along -z after rotation: M = D R T
    M = b.reshape(4, 4).T
    assert np.allclose(M[:3, :3], R) and abs(M[2, 3] - (-3.3 + (R @ [0.1, -0.2, 0.3])[2])) < 1e-12


def _inv(oracle_lib, m, f32):
    if f32:
        a = (ctypes.c_float * 16)(*m.T.reshape(16))
        o = (ctypes.c_float * 16)()
        oracle_lib.lib().oracle_inverse_f32(a, o)
    else:
        a = (ctypes.c_double * 16)(*m.T.reshape(16))
        o = (ctypes.c_double * 16)()
        oracle_lib.lib().oracle_inverse_f64(a, o)
    return np.array(o).reshape(4, 4).T


def test_inverse(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(50):
        m = rng.normal(size=(4, 4)) + 4 * np.eye(4)
        i64 = _inv(oracle_lib, m, False)
        assert np.allclose(i64 @ m, np.eye(4), atol=1e-12)
        i32 = _inv(oracle_lib, m.astype(np.float32).astype(np.float64), True)
        assert np.allclose(i32, np.linalg.inv(m.astype(np.float32).astype(np.float64)), rtol=1e-4, atol=1e-5)
    d = np.diag([2.0, 4.0, 8.0, 16.0])
    assert np.array_equal(_inv(oracle_lib, d, True), np.diag([0.5, 0.25, 0.125, 0.0625]))
    t = np.eye(4)
    t[:3, 3] = (1.5, -2.0, 4.0)
    it = np.eye(4)
    it[:3, 3] = (-1.5, 2.0, -4.0)
    assert np.array_equal(_inv(oracle_lib, t, True), it)


# --------------------------------------------------------------------------- rays
def _textbook_ray(ps, px, py, clip=()):
    """pixel -> ray in float64 with numpy's inverse and a slab test against [0, L]^3."""
    spec = ps.spec
    W, H = spec.framebuffer
    P = ps.proj.reshape(4, 4).T
    MV = ps.mv.reshape(4, 4).T
    inv = np.linalg.inv(P @ MV)
    ndc = (px / W * 2 - 1, py / H * 2 - 1)
    s = inv @ np.array([ndc[0], ndc[1], -1, 1.0])
    e = inv @ np.array([ndc[0], ndc[1], 1, 1.0])
    gs = np.array(spec.global_size, np.float64) * np.array(spec.scale)
    half = max(np.floor(gs)) * 0.5
    start = s[:3] / s[3] * half
    end = e[:3] / e[3] * half
    d = (end - start) / np.linalg.norm(end - start)
    start = start + np.floor(np.floor(gs) / 2)
    L = np.floor(np.array(spec.local_size) * np.array(spec.scale))
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (0 - start) / d
        t1 = (L - start) / d
    tmin, tmax = np.minimum(t0, t1), np.maximum(t0, t1)
    enter, leave = tmin.max(), tmax.min()
    for pos, n in clip:
        n = np.asarray(n, np.float64) * spec.scale
        n = n / np.linalg.norm(n)
        p = np.asarray(pos) * np.floor(gs) * 0.5 + np.floor(np.floor(gs) / 2)
        dn = d @ n
        t = ((p - start) @ n) / dn
        if dn > 0:
            if leave < t:
                return None
            enter = max(enter, t)
        else:
            if enter > t:
                return None
            leave = min(leave, t)
    enter = max(enter, float(np.float32(0.002)))
    if enter > leave:
        return None
    return d, start, enter, leave


def _oracle_ray(oracle_lib, ps, precision, px, py):
    out = (ctypes.c_double * 14)()
    hit = oracle_lib.lib().oracle_ray(ctypes.byref(ps.sc), precision, px, py, out)
    return hit, np.array(out)


@pytest.mark.parametrize("clip", [[], [((0.0, 0.0, 0.0), (1.0, 0.0, 0.0))], [((0.2, -0.1, 0.0), (1.0, 1.0, -0.5))]])
def test_rays_against_textbook_slab(oracle_lib, clip):
    spec = dataclasses.replace(scenes.small_scene(n=32, fb=(64, 48)), clipping=clip, camera=scenes.FILL_POSE)
    ps = oracle_lib.PreparedScene(spec)
    agree = total = 0
    for py in range(0, 48, 3):
        for px in range(0, 64, 3):
            tb = _textbook_ray(ps, px, py, clip)
            h64, r64 = _oracle_ray(oracle_lib, ps, 64, px, py)
            h32, r32 = _oracle_ray(oracle_lib, ps, 32, px, py)
            total += 1
            if tb is None:
                agree += (h64 == 0)
                continue
            d, start, enter, leave = tb
            if not h64:
                continue
            agree += 1
            assert np.allclose(r64[0:3], d, atol=1e-12)
            assert np.allclose(r64[3:6], start, rtol=1e-10, atol=1e-9)
            assert abs(r64[6] - enter) < 1e-8 * max(1, abs(enter)) and abs(r64[7] - leave) < 1e-8 * max(1, abs(leave))
            if h32:
                # binary32 inverse of P*MV (near plane 0.002) limits the direction to ~1e-5
                assert np.allclose(r32[0:3], d, atol=5e-5)
                assert abs(r32[6] - enter) < 2e-3 * max(1, abs(enter))
    assert agree >= total - 2


def test_outer_face_clip_normal(oracle_lib):
    spec = scenes.small_scene(n=32, fb=(64, 64))
    ps = oracle_lib.PreparedScene(spec)
    hit, r = _oracle_ray(oracle_lib, ps, 64, 32, 32)
    assert hit and r[8] == 1.0
    # the default camera looks down -z and enters through the z = max face: normal z = sign(dir.z) = -1
    assert r[2] < 0 and tuple(r[9:12]) == (0.0, 0.0, -1.0)


# --------------------------------------------------------------------------- whole-pass closed forms
def _const_scene(alpha, n=24, fb=(96, 72), interpolation=False, pose=scenes.DEFAULT_POSE):
    data = np.full((n, n, n), 0.6, np.float32)
    col = (0.8, 0.4, 0.2, alpha)
    vol = scenes.VolumeSourceSpec("c", data, "idem", [(0, col), (1024, col)], 2.0, 0.0)
    return scenes.SceneSpec(
        name="const", global_size=(n,) * 3, local_size=(n,) * 3, particle_grid=(4, 4, 4), framebuffer=fb,
        volumes=[vol], interpolation=interpolation, dither_mode=0, camera=pose,
    )


@pytest.mark.parametrize("alpha", [0.02, 0.9])
def test_volume_compositing_closed_form(oracle_lib, alpha):
    """Constant premultiplied texel v: C_k = v_rgb * (1-(1-a)^k)/a, C.a_k = 1-(1-a)^k, early exit at 0.99."""
    spec = _const_scene(alpha)
    res = oracle_lib.render(spec, 64)
    ps = res.prepared
    texel = np.round(np.clip(np.array([0.8 * alpha, 0.4 * alpha, 0.2 * alpha, alpha]), 0, 1) * 255) / 255
    W, H = spec.framebuffer
    total_steps = 0
    ok = n = 0
    for py in range(H):
        for px in range(W):
            tb = _textbook_ray(ps, px, py)
            if tb is None:
                continue
            d, start, enter, leave = tb
            s = spec.step
            factor = s / spec.global_size[0] * 2 * 2.0
            a = texel[3] * factor
            k = math.floor(leave / s) - math.ceil(enter / s) + 1
            if k <= 0:
                continue
            kexit = math.ceil(math.log(0.01) / math.log(1 - a)) if a > 0 else k
            kk = min(k, kexit)
            # the loop stops at the first k with C.a > 0.99
            while kk > 1 and 1 - (1 - a) ** (kk - 1) > 0.99:
                kk -= 1
            total_steps += kk
            frac = 1 - (1 - a) ** kk
            c = np.append(texel[:3] * factor * frac / a, frac)
            exp = np.floor(np.clip(c, 0, 1) * 255)
            n += 1
            ok += np.all(np.abs(res.rgba[py, px].astype(float) - exp) <= 1)
    assert n > 500 and ok >= 0.995 * n, (ok, n)
    assert abs(res.counters["vol_samples"] - total_steps) <= 0.002 * total_steps


def test_iso_plane_depth_and_normal(oracle_lib):
    """Field 1 - z/Z with an alpha ramp TF and iso threshold 0.5: the surface is the plane z = Z/2."""
    n = 32
    z = (np.arange(n) + 0.5) / n
    data = np.broadcast_to((1.0 - z)[:, None, None], (n, n, n)).astype(np.float32).copy()
    vol = scenes.VolumeSourceSpec("ramp", data, "idem", [(0, (0.2, 0.6, 1.0, 0.0)), (1024, (0.2, 0.6, 1.0, 1.0))], 0.0, 0.5)
    spec = scenes.SceneSpec(
        name="iso", global_size=(n,) * 3, local_size=(n,) * 3, particle_grid=(4, 4, 4), framebuffer=(80, 60),
        volumes=[vol], interpolation=True, step=0.25,
    )
    res = oracle_lib.render(spec, 64)
    ps = res.prepared
    checked = 0
    for py in range(0, 60, 2):
        for px in range(0, 80, 2):
            tb = _textbook_ray(ps, px, py)
            if tb is None:
                continue
            d, start, enter, leave = tb
            t = (n / 2 - start[2]) / d[2]
            hitp = start + t * d
            if not (3 < hitp[0] < n - 3 and 3 < hitp[1] < n - 3) or t < enter + 1:
                continue
            assert abs(res.pre_depth[py, px] - t) < 0.6, (px, py, res.pre_depth[py, px], t)
            nrm = res.normal[py, px]
            assert nrm[2] > 0.999, nrm
            checked += 1
    assert checked > 60
    assert res.counters["iso_hits"] >= checked


def _particle_scene(pos, rad=0.4, fb=(96, 72), pose=scenes.FILL_POSE, n=16, grid=4):
    data = np.zeros((n, n, n), np.float32)
    vol = scenes.VolumeSourceSpec("empty", data, "idem", [(0, (0, 0, 0, 0)), (1024, (0, 0, 0, 0))], 0.0, 0.0)
    attr = np.linspace(0.2, 0.9, len(pos), dtype=np.float32)[:, None]
    part = scenes.ParticleSourceSpec("p", np.asarray(pos, np.float32), attr, None, rad, "idem",
                                     [(0, (1.0, 0.0, 0.0, 1.0)), (1024, (0.0, 0.0, 1.0, 1.0))], 2.0)
    return scenes.SceneSpec(
        name="parts", global_size=(n,) * 3, local_size=(n,) * 3, particle_grid=(grid,) * 3, framebuffer=fb,
        volumes=[vol], particles=[part], interpolation=True, camera=pose,
    )


def test_particles_brute_force_nearest_sphere(oracle_lib):
    rng = np.random.default_rng(7)
    cells = rng.choice(64, size=30, replace=False)
    cc = np.stack([cells % 4, (cells // 4) % 4, cells // 16], 1)
    pos = cc * 4 + 2 + rng.uniform(-1, 1, size=(30, 3))  # sphere radius 0.8 stays inside its 4-voxel cell
    spec = _particle_scene(pos)
    res = oracle_lib.render(spec, 64)
    ps = res.prepared
    W, H = spec.framebuffer
    radius = 0.4 * 2.0
    checked = hits = 0
    for py in range(H):
        for px in range(W):
            tb = _textbook_ray(ps, px, py)
            if tb is None:
                continue
            d, start, enter, leave = tb
            best = None
            for i, p in enumerate(pos):
                L = p - start
                tca = L @ d
                d2 = L @ L - tca * tca
                if d2 > radius * radius:
                    continue
                thc = math.sqrt(radius * radius - d2)
                t0, t1 = tca - thc, tca + thc
                if t1 >= enter and (best is None or t0 < best[0]):
                    best = (t0, i)
            depth = res.pre_depth[py, px]
            checked += 1
            if best is None:
                assert depth > 1e30, (px, py)
            else:
                hits += 1
                assert abs(depth - best[0]) < 1e-4, (px, py, depth, best)
    assert hits > 50 and checked > 1000


def test_particles_first_hit_cell_rule(oracle_lib):
    """A big sphere reaching back into an earlier cell loses against a small sphere
    inside that earlier cell (lib/isaac/isaac_particle_kernel.hpp:204-205)."""
    near = np.array([[8.0, 8.0, 12.5]])  # in cell z = 3 ([12,16)), visited first: the camera looks down -z
    far = np.array([[8.0, 8.0, 10.0]])  # in cell z = 2 ([8,12)); radius 3.5 reaches z = 13.5 inside cell 3
    pos = np.concatenate([far, near])
    spec = _particle_scene(pos, rad=0.25, fb=(64, 64), pose=scenes.DEFAULT_POSE)
    p = spec.particles[0]
    radius = np.array([3.5 / 2.0, 0.2 / 2.0], np.float32)  # times weight 2 -> 3.5 and 0.2
    spec.particles[0] = dataclasses.replace(p, radius=radius)
    res = oracle_lib.render(spec, 64)
    ps = res.prepared
    d, start, enter, leave = _textbook_ray(ps, 32, 32)
    assert d[2] < -0.99
    depth = res.pre_depth[32, 32]
    # along the centre ray: the big sphere's surface (z = 13.5) comes before the small one's (z = 12.7),
    # but the DDA stops in the first cell that holds a hit, cell 3, whose list has only the small sphere
    t_small = (start[2] - 12.7) / -d[2]
    t_big = (start[2] - 13.5) / -d[2]
    assert t_big < t_small
    assert abs(depth - t_small) < 0.05, (depth, t_small, t_big)


def test_binning_brute_force(oracle_lib):
    rng = np.random.default_rng(11)
    pos = rng.uniform(0, 32, size=(5000, 3)).astype(np.float32)
    pos[:10] = [[0, 0, 0], [4, 4, 4], [31.999, 0, 8], [8, 8, 8], [7.9999995, 16, 24], [32, 32, 32],
                [-0.5, 3, 3], [12, 12, 12], [28, 28, 28], [3.9999998, 4.0000005, 0]]
    grid = (8, 8, 8)
    cell_of, counts, offsets, perm, rel = oracle_lib.bin_particles(pos, (32, 32, 32), grid)
    cw = np.float32(32) / np.float32(8)
    q = pos / cw
    c = np.clip(np.floor(q).astype(np.int64), 0, 7)
    cid = c[:, 0] + 8 * (c[:, 1] + 8 * c[:, 2])
    assert np.array_equal(cell_of, cid)
    assert np.array_equal(counts, np.bincount(cid, minlength=512))
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(np.bincount(cid, minlength=512))]))
    order = np.argsort(cid, kind="stable")
    assert np.array_equal(perm, order)
    exp_rel = (q - c.astype(np.float32)).astype(np.float32)[order]
    assert np.array_equal(rel, exp_rel)


def test_ssao_brute_force(oracle_lib):
    spec = dataclasses.replace(scenes.prod_scene(n=32, fb=(90, 70), n_particles=3000), ao_enabled=True)
    p = spec.particles[0]
    spec.particles[0] = dataclasses.replace(p, tf_points=[(0, (1, 1, 1, 1)), (1024, (1, 1, 1, 1))])
    res = oracle_lib.render(spec, 32)
    dep = res.pre_depth
    H, W = dep.shape
    exp = np.zeros_like(dep)
    for y in range(H):
        for x in range(W):
            cnt = 0
            for i in range(-3, 4):
                for j in range(-3, 4):
                    sx = min(max(x + 10 * i, 0), W - 1)
                    sy = min(max(y + 10 * j, 0), H - 1)
                    cnt += dep[sy, sx] < dep[y, x]
            exp[y, x] = min(max(np.float32(cnt) / np.float32(42), 0), 1)
    assert np.array_equal(res.ao, exp)
    assert (res.ao > 0).any()


@pytest.mark.parametrize("mode", [0, 1, 3, 4, 6])
def test_shading_closed_form(oracle_lib, mode):
    spec = dataclasses.replace(scenes.prod_scene(n=32, fb=(64, 48), n_particles=3000), render_mode=mode, ao_weight=0.7)
    p = spec.particles[0]
    spec.particles[0] = dataclasses.replace(p, tf_points=[(0, (0.3, 0.6, 0.9, 1)), (1024, (0.9, 0.2, 0.4, 1))])
    res = oracle_lib.render(spec, 64)
    ps = res.prepared
    H, W = res.pre_depth.shape
    P = ps.proj.reshape(4, 4).T
    MV = ps.mv.reshape(4, 4).T
    inv = np.linalg.inv(P @ MV)
    bad = 0
    for y in range(H):
        for x in range(W):
            c = res.pre[y, x].astype(np.float64) / 255
            nrm = res.normal[y, x].astype(np.float64)
            ao = float(res.ao[y, x])
            ndc = (x / W * 2 - 1, y / H * 2 - 1)
            s = inv @ np.array([ndc[0], ndc[1], -1, 1.0])
            e = inv @ np.array([ndc[0], ndc[1], 1, 1.0])
            d = e[:3] / e[3] - s[:3] / s[3]
            d /= np.linalg.norm(d)
            w = 0.7
            aof = (1 - w) + w * (1 - ao)
            if mode in (0, 1):
                lf = (abs(nrm @ -d) + aof) * 0.5
                spec_ = 0.0 if mode == 1 else (nrm @ -d) ** 4 * 0.5
                exp = np.append(np.minimum(c[:3] * lf + spec_, 1), c[3])
            elif mode == 3:
                exp = np.array([0, 0, 0, c[3]])
            elif mode == 4:
                exp = np.append(nrm * 0.5 + 0.5, c[3])
            else:
                exp = np.array([aof, aof, aof, c[3]])
            exp = np.floor(np.clip(exp, 0, 1) * 255)
            bad += np.any(np.abs(res.solid[y, x].astype(float) - exp) > 1)
    assert bad <= 2, bad


def test_fp32_oracle_close_to_fp64(oracle_lib):
    spec = scenes.small_scene()
    a = oracle_lib.render(spec, 32)
    b = oracle_lib.render(spec, 64)
    p = oracle_lib.psnr(a.rgba, b.rgba)
    print(f"small64 oracle FP32 vs FP64: PSNR {p:.2f} dB")
    assert p > 60
    d = np.abs(a.rgba.astype(int) - b.rgba.astype(int)).max(axis=2)
    assert (d <= 2).mean() >= 0.999
