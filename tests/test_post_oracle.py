"""Pins of the oracle's brick (post-classification) form, all on the CPU.

Each check uses something other than the oracle's own formulas: numpy's float16
and torch's bfloat16 casts for the 16-bit encodings, a numpy trilinear sampler
and brick encoder, closed forms for compositing through a constant field and for
an iso-plane, and the pre-classification oracle (itself pinned in
test_oracle_kats.py) for the shared march structure.
"""
import dataclasses
import math

import numpy as np
import pytest

from isaac_amd import scenes

from test_oracle_kats import _const_scene, _textbook_ray

RAW, F32, F16, BF16 = -1, 0, 1, 2


# --------------------------------------------------------------------------- encodings
def _probe_values():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(300_000).astype(np.float32) * (np.float32(10.0) ** rng.integers(-10, 6, 300_000)).astype(np.float32)
    edges = np.array(
        [0.0, -0.0, 1.0, 65504.0, 65519.99, 65520.0, 1e9, 6.1035156e-05, 6.0e-05, 5.9604645e-08, 2.9802322e-08,
         2.9802326e-08, 1.0e-8, np.inf, -np.inf, 1.0009766, 1.0004883, 1.0014648, 3.0e38],
        np.float32,
    )
    return np.concatenate([x, edges, -edges])


def test_f16_rne_matches_numpy(oracle_lib):
    x = _probe_values()
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(oracle_lib.f32_to_f16_bits(x), want)
    # decode of every finite half
    h = np.arange(0, 1 << 16, dtype=np.uint16)
    ref = h.view(np.float16).astype(np.float32)
    got = oracle_lib.f16_bits_to_f32(h)
    same = (got.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(got) & np.isnan(ref))
    assert same.all()


def test_bf16_rne_matches_torch(oracle_lib):
    import torch

    x = _probe_values()
    want = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(oracle_lib.f32_to_bf16_bits(x), want)


def _brick_scales_numpy(data):
    if data.ndim == 3:
        data = data[..., None]
    Z, Y, X, C = data.shape
    out = []
    for bz in range((Z + 7) // 8):
        for by in range((Y + 7) // 8):
            for bx in range((X + 7) // 8):
                zi = np.minimum(bz * 8 + np.arange(9), Z - 1)
                yi = np.minimum(by * 8 + np.arange(9), Y - 1)
                xi = np.minimum(bx * 8 + np.arange(9), X - 1)
                r = np.abs(data[np.ix_(zi, yi, xi)]).reshape(-1, C).max(axis=0)
                out.append(np.where(r > 0, r, 1.0))
    return np.array(out, np.float32)


@pytest.mark.parametrize("shape", [(16, 16, 16), (13, 9, 21), (8, 17, 3, 3)])
def test_brick_scales_brute_force(oracle_lib, shape):
    rng = np.random.default_rng(3)
    data = rng.standard_normal(shape).astype(np.float32)
    data[:8, :8, :8] = 0.0  # an all-zero brick region takes scale 1 ... unless its apron is not zero
    want = _brick_scales_numpy(data)
    assert np.array_equal(oracle_lib.brick_scales(data, F16), want)
    assert np.array_equal(oracle_lib.brick_scales(data, BF16), want)
    assert np.all(oracle_lib.brick_scales(data, F32) == 1.0)


def test_zero_brick_scale_is_one(oracle_lib):
    data = np.zeros((20, 20, 20), np.float32)
    data[19, 19, 19] = 2.0
    s = oracle_lib.brick_scales(data, F16)
    assert s[0, 0] == 1.0 and s[-1, 0] == 2.0


# --------------------------------------------------------------------------- sampler
def _numpy_trilinear(data, p, quant=None):
    """Sampler<LINEAR, CLAMP>: taps at floor(p - 0.5) and +1, clamped; z innermost."""
    Z, Y, X = data.shape[:3]
    c = p - 0.5
    f = np.floor(c)
    w = c - f
    b = f.astype(np.int64)
    x0, y0, z0 = np.clip(b[0], 0, X - 1), np.clip(b[1], 0, Y - 1), np.clip(b[2], 0, Z - 1)
    x1, y1, z1 = np.clip(b[0] + 1, 0, X - 1), np.clip(b[1] + 1, 0, Y - 1), np.clip(b[2] + 1, 0, Z - 1)
    d = lambda x, y, z: float(data[z, y, x]) if quant is None else quant(x0, y0, z0, float(data[z, y, x]))
    a0 = d(x0, y0, z0) * (1 - w[2]) + d(x0, y0, z1) * w[2]
    a1 = d(x0, y1, z0) * (1 - w[2]) + d(x0, y1, z1) * w[2]
    b0 = d(x1, y0, z0) * (1 - w[2]) + d(x1, y0, z1) * w[2]
    b1 = d(x1, y1, z0) * (1 - w[2]) + d(x1, y1, z1) * w[2]
    A = a0 * (1 - w[1]) + a1 * w[1]
    B = b0 * (1 - w[1]) + b1 * w[1]
    return A * (1 - w[0]) + B * w[0], (x0, y0, z0)


def test_raw_sampler_against_numpy(oracle_lib):
    rng = np.random.default_rng(11)
    data = rng.random((12, 10, 14)).astype(np.float32)
    pos = rng.uniform(-1.5, 15.5, size=(400, 3))
    got = oracle_lib.sample_field(data, RAW, pos, linear=True, precision=64)[:, 0]
    want = np.array([_numpy_trilinear(data, p)[0] for p in pos])
    assert np.allclose(got, want, rtol=0, atol=1e-12)
    near = oracle_lib.sample_field(data, RAW, pos, linear=False, precision=64)[:, 0]
    idx = np.clip(np.trunc(pos).astype(np.int64), 0, [13, 9, 11])
    assert np.array_equal(near, data[idx[:, 2], idx[:, 1], idx[:, 0]].astype(np.float64))


@pytest.mark.parametrize("dtype", [F16, BF16])
def test_brick_sampler_quantises_per_brick(oracle_lib, dtype):
    """16-bit bricks: each tap is RN16(v / s_B) * s_B with B the brick of the sample's
    (x0, y0, z0) tap -- apron taps take that brick's scale, not their own."""
    import torch

    rng = np.random.default_rng(5)
    data = (rng.random((20, 18, 22)) * np.linspace(0.01, 10.0, 22)[None, None, :]).astype(np.float32)
    scales = oracle_lib.brick_scales(data, dtype)[:, 0]
    nbx, nby = (22 + 7) // 8, (18 + 7) // 8

    def enc(x):
        x = np.float32(x)
        if dtype == F16:
            return float(np.float32(np.float16(x)))
        return float(torch.tensor([x]).to(torch.bfloat16).float().item())

    def quant(x0, y0, z0, v):
        s = scales[((z0 // 8) * nby + y0 // 8) * nbx + x0 // 8]
        return enc(np.float32(v) / s)

    pos = rng.uniform(-1.0, 23.0, size=(500, 3))
    pos[:100, 0] = 8.0 + rng.uniform(-0.5, 0.5, 100)  # straddle the x = 8 brick face
    got = oracle_lib.sample_field(data, dtype, pos, linear=True, precision=64)[:, 0]
    want = []
    for p in pos:
        v, (x0, y0, z0) = _numpy_trilinear(data, p, quant)
        want.append(v * scales[((z0 // 8) * nby + y0 // 8) * nbx + x0 // 8])
    assert np.allclose(got, np.array(want), rtol=1e-12, atol=1e-12)
    # and the encoding error stays within the format's half ulp of the brick maximum
    raw = oracle_lib.sample_field(data, RAW, pos, linear=True, precision=64)[:, 0]
    ulp = 2.0 ** -11 if dtype == F16 else 2.0 ** -8
    assert np.all(np.abs(got - raw) <= ulp * np.max(data) + 1e-12)


# --------------------------------------------------------------------------- whole-frame pins
@pytest.mark.parametrize("alpha", [0.02, 0.9])
@pytest.mark.parametrize("dtype", [RAW, F32, F16])
def test_post_compositing_closed_form(oracle_lib, alpha, dtype):
    """Constant field: every sample classifies to the TF colour (no RGBA8 texel), so
    C_k = v_rgb (1-(1-a)^k)/a, C.a_k = 1-(1-a)^k with a = alpha * factor."""
    spec = _const_scene(alpha)
    res = oracle_lib.render_post(spec, 64, dtype)
    ps = res.prepared
    v = np.array([0.8 * alpha, 0.4 * alpha, 0.2 * alpha, alpha])
    W, H = spec.framebuffer
    total = ok = n = 0
    for py in range(H):
        for px in range(W):
            tb = _textbook_ray(ps, px, py)
            if tb is None:
                continue
            d, start, enter, leave = tb
            s = spec.step
            factor = s / spec.global_size[0] * 2 * 2.0
            a = v[3] * factor
            k = math.floor(leave / s) - math.ceil(enter / s) + 1
            if k <= 0:
                continue
            kk = k
            for j in range(1, k + 1):
                if 1 - (1 - a) ** j > 0.99:
                    kk = j
                    break
            total += kk
            frac = 1 - (1 - a) ** kk
            c = np.append(v[:3] * factor * frac / a, frac)
            n += 1
            # the TF table holds binary32 colours (one rounding of each channel)
            ok += np.all(np.abs(res.rgba_f[py, px] - c) <= 2e-6)
    assert n > 500 and ok >= 0.995 * n, (ok, n)
    assert abs(res.counters["vol_samples"] - total) <= 0.002 * total


def test_post_iso_plane(oracle_lib):
    """Linear field 1 - z/Z, alpha ramp TF, threshold 0.5: post-classification puts the
    surface on the plane z = Z/2 (no RGBA8 texel rounding in the way)."""
    n = 32
    z = (np.arange(n) + 0.5) / n
    data = np.broadcast_to((1.0 - z)[:, None, None], (n, n, n)).astype(np.float32).copy()
    vol = scenes.VolumeSourceSpec("ramp", data, "idem", [(0, (0.2, 0.6, 1.0, 0.0)), (1024, (0.2, 0.6, 1.0, 1.0))], 0.0, 0.5)
    spec = scenes.SceneSpec(
        name="iso", global_size=(n,) * 3, local_size=(n,) * 3, particle_grid=(4, 4, 4), framebuffer=(80, 60),
        volumes=[vol], interpolation=True, step=0.25,
    )
    res = oracle_lib.render_post(spec, 64, RAW)
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
            # the reference interpolates forward from step i (t_i + s (0.5 - old) / (v - old),
            # lib/isaac/isaac_iso_kernel.hpp:84-92), so the depth lies up to one step past the plane
            e = res.depth[py, px] - t
            assert -0.05 < e < spec.step + 0.05, (px, py, res.depth[py, px], t)
            assert res.normal[py, px][2] > 0.999
            checked += 1
    assert checked > 60 and res.counters["iso_hits"] >= checked


@pytest.mark.parametrize("interpolation", [True, False])
def test_post_agrees_with_pre_classification(oracle_lib, interpolation):
    """Same march structure as the pre-classification oracle: equal ray counts, sample
    counts within early-exit jitter, images close (the difference is the RGBA8 texel)."""
    spec = dataclasses.replace(scenes.small_scene(n=48, fb=(160, 120)), interpolation=interpolation)
    pre = oracle_lib.render(spec, 64)
    post = oracle_lib.render_post(spec, 64, RAW)
    assert post.counters["vol_rays"] == pre.counters["vol_rays"]
    assert abs(post.counters["vol_samples"] - pre.counters["vol_samples"]) <= 2e-3 * pre.counters["vol_samples"]
    assert oracle_lib.psnr(post.rgba, pre.rgba) > 40.0


def test_post_production_passes(oracle_lib):
    """Particles, iso, clipping and SSAO in the brick form reuse the pre-classification
    passes: the particle walk counts are identical, iso counts close."""
    spec = scenes.prod_scene(n=48, fb=(240, 135), n_particles=8000)
    pre = oracle_lib.render(spec, 64)
    post = oracle_lib.render_post(spec, 64, RAW)
    for k in ("cells_visited", "sphere_tests", "part_rays", "iso_rays"):
        assert post.counters[k] == pre.counters[k], k
    assert abs(post.counters["iso_samples"] - pre.counters["iso_samples"]) <= 0.02 * pre.counters["iso_samples"]
    assert oracle_lib.psnr(post.rgba, pre.rgba) > 30.0


@pytest.mark.parametrize("dtype", [F32, F16, BF16])
def test_post_fp32_close_to_fp64(oracle_lib, dtype):
    spec = scenes.small_scene()
    r64 = oracle_lib.render_post(spec, 64, RAW)
    r = oracle_lib.render_post(spec, 32, dtype)
    floor = {F32: 100.0, F16: 80.0, BF16: 70.0}[dtype]
    assert oracle_lib.psnr_float(r.rgba_f, r64.rgba_f) > floor


def test_sampled_rendering_matches_full(oracle_lib):
    spec = scenes.small_scene(n=32, fb=(96, 64))
    full = oracle_lib.render_post(spec, 64, RAW)
    part = oracle_lib.render_post(spec, 64, RAW, stride=4)
    assert np.array_equal(full.rgba_f[::4, ::4], part.rgba_f[::4, ::4])


@pytest.mark.parametrize("stride", [2, 5, 10])
def test_sampled_rendering_with_ssao_matches_full(oracle_lib, stride):
    """With SSAO on, the sampled pixel set is closed under the 7x7 taps (spacing 10,
    clamped at the borders), so every sampled pixel equals the full render's."""
    import dataclasses

    spec = scenes.prod_scene(n=32, fb=(201, 133), n_particles=500, pose=scenes.FILL_POSE)
    spec = dataclasses.replace(spec, ao_enabled=True, ao_weight=0.8)
    full = oracle_lib.render_post(spec, 32, F16)
    part = oracle_lib.render_post(spec, 32, F16, stride=stride)
    m = oracle_lib.sampled_mask(201, 133, stride, True)
    assert m[0].sum() > 1 and m[:, 0].sum() > 1 and m[-1, -1]
    assert (full.ao[m] > 0).any(), "the scene should occlude somewhere"
    assert np.array_equal(full.ao[m], part.ao[m])
    assert np.array_equal(full.rgba[m], part.rgba[m])
    assert np.array_equal(full.rgba_f[m], part.rgba_f[m])


def test_sampled_rendering_with_ssao_rejects_bad_strides(oracle_lib):
    import dataclasses

    spec = dataclasses.replace(scenes.small_scene(n=16, fb=(40, 30)), ao_enabled=True)
    with pytest.raises(RuntimeError):
        oracle_lib.render_post(spec, 32, F16, stride=3)
