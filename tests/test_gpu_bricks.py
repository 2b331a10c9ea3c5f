"""GPU parity of the brick render form (FP32 / FP16 / BF16 field bricks,
post-classification march).

Bars:
  * brick contents (pairs and scales) ............ bit-exact against an encoder
    written here in numpy (float16 casts) and torch (bfloat16 casts)
  * RGBA8 framebuffer vs the oracle's brick form in the library's arithmetic
    (binary64 ray setup, binary32 sampling) ...... |diff| <= 2/255 on >= 99.9 % of pixels
  * float image vs the FP64 oracle (raw float sources, everything binary64):
    PSNR >= 40 dB (peak 1), stated with max / mean absolute error
"""
import dataclasses

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 2
FRACTION = 0.999
PSNR_FLOOR_DB = 40.0
F32, F16, BF16 = 0, 1, 2
NAMES = {F32: "f32", F16: "f16", BF16: "bf16"}


def brick_frame(spec, gpu, dtype, counting=True):
    from isaac_amd import visualization as V

    vis = V.from_scene(spec, gpu.index or 0)
    vis.set_render_form(vis.FORM_BRICKS, dtype)
    vis.set_float_output(True)
    vis.set_counting(counting)
    img = vis.do_visualization()
    vis.synchronize()
    return vis, img, vis.float_image()


def expected_bricks(data, dtype):
    """Pairs (nbricks, 9, 9, 8, dim, 2) and scales (nbricks, dim) of a (Z, Y, X[, C]) source."""
    import torch

    if data.ndim == 3:
        data = data[..., None]
    Z, Y, X, C = data.shape
    nbx, nby, nbz = (X + 7) // 8, (Y + 7) // 8, (Z + 7) // 8
    bz, by, bx = np.meshgrid(np.arange(nbz), np.arange(nby), np.arange(nbx), indexing="ij")
    bz, by, bx = bz.reshape(-1), by.reshape(-1), bx.reshape(-1)
    l9 = np.arange(9)
    zi = np.minimum(bz[:, None] * 8 + l9[None, :], Z - 1)  # (nb, 9)
    yi = np.minimum(by[:, None] * 8 + l9[None, :], Y - 1)
    xi = np.minimum(bx[:, None] * 8 + l9[None, :], X - 1)
    region = data[zi[:, :, None, None], yi[:, None, :, None], xi[:, None, None, :], :]  # (nb, 9, 9, 9, C)
    if dtype == F32:
        scales = np.ones((region.shape[0], C), np.float32)
    else:
        m = np.abs(region).reshape(region.shape[0], -1, C).max(axis=1)
        scales = np.where(m > 0, m, 1.0).astype(np.float32)
    v0 = region[:, :, :, 0:8, :]
    v1 = region[:, :, :, 1:9, :]
    pairs = np.stack([v0, v1], axis=-1)  # (nb, 9z, 9y, 8x, C, 2)
    if dtype == F32:
        return pairs.astype(np.float32), scales
    q = (pairs / scales[:, None, None, None, :, None]).astype(np.float32)
    if dtype == F16:
        return q.astype(np.float16).view(np.uint16), scales
    t = torch.from_numpy(np.ascontiguousarray(q)).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    return t, scales


@pytest.mark.parametrize("dtype", [F32, F16, BF16])
def test_brick_contents_bit_exact(gpu, oracle_lib, dtype):
    """Odd-sized box (edge bricks clamp), scalar + 3-component sources."""
    from oracle import oracle

    spec = scenes.prod_scene(n=44, fb=(64, 48), n_particles=1000)
    vis, _, _ = brick_frame(spec, gpu, dtype, counting=False)
    for s, v in enumerate(spec.volumes):
        pairs, scales = vis.bricks(s)
        want_pairs, want_scales = expected_bricks(v.data, dtype)
        assert np.array_equal(scales, want_scales), f"source {s} {NAMES[dtype]} scales"
        assert np.array_equal(scales, oracle.brick_scales(v.data, dtype)), "oracle brick scales"
        if dtype == F32:
            assert np.array_equal(pairs.view(np.uint32), want_pairs.view(np.uint32)), f"source {s} pairs"
        else:
            assert np.array_equal(pairs, want_pairs), f"source {s} {NAMES[dtype]} pairs"
    vis.close()


def compare_post(label, spec, img, fimg, vis, dtype, ora, ora64):
    from oracle import oracle

    d = np.abs(img.astype(np.int32) - ora.rgba.astype(np.int32))
    ok = (d.max(axis=2) <= TOL).mean()
    fd = np.abs(fimg.astype(np.float64) - ora.rgba_f.astype(np.float64))
    e64 = np.abs(fimg.astype(np.float64) - ora64.rgba_f.astype(np.float64))
    p64 = oracle.psnr_float(fimg, ora64.rgba_f)
    print(
        f"{label} {NAMES[dtype]}: within {TOL}/255 of the oracle on {ok * 100:.4f}% (max {d.max()}), "
        f"float max diff {fd.max():.3g}; vs FP64: PSNR {p64:.2f} dB, max abs {e64.max():.4g}, "
        f"mean abs {e64.mean():.3g}",
        flush=True,
    )
    assert ok >= FRACTION
    c = vis.counters()
    for k in ("vol_samples", "iso_samples", "iso_hits", "cells_visited", "vol_rays", "iso_rays", "part_rays"):
        rel = abs(c[k] - ora.counters[k]) / max(1, ora.counters[k])
        assert rel <= 1e-3, (k, c[k], ora.counters[k])
    return p64


@pytest.mark.parametrize("dtype", [F32, F16, BF16])
def test_small_scene_bricks(gpu, oracle_lib, dtype):
    from oracle import oracle

    spec = scenes.small_scene()
    vis, img, fimg = brick_frame(spec, gpu, dtype)
    ora = oracle.render_post(spec, 32, dtype)
    ora64 = oracle.render_post(spec, 64, oracle.BRICK_RAW)
    p = compare_post("small64", spec, img, fimg, vis, dtype, ora, ora64)
    assert p >= PSNR_FLOOR_DB
    vis.close()


@pytest.mark.parametrize("dtype", [F32, F16, BF16])
@pytest.mark.parametrize("pose", ["default", "fill"])
def test_reduced_production_bricks(gpu, oracle_lib, dtype, pose):
    """3 sources (density, curl field via length, particles), iso, clipping, SSAO."""
    from oracle import oracle

    spec = scenes.prod_scene(n=64, fb=(480, 270), n_particles=20000,
                             pose=scenes.DEFAULT_POSE if pose == "default" else scenes.FILL_POSE)
    vis, img, fimg = brick_frame(spec, gpu, dtype)
    ora = oracle.render_post(spec, 32, dtype)
    ora64 = oracle.render_post(spec, 64, oracle.BRICK_RAW)
    p = compare_post(f"prod64-{pose}", spec, img, fimg, vis, dtype, ora, ora64)
    color, depth, normal, ao = vis.gbuffer()
    both = (ora.depth < 3e38) & (depth < 3e38)
    assert (np.abs(depth[both] - ora.depth[both]) <= 1e-3 * np.maximum(1.0, np.abs(ora.depth[both]))).mean() >= FRACTION
    # iso edges flip between precisions; the image stays above the floor
    assert p >= 30.0
    # the march compiled for the source signature (natural chains inlined) renders the
    # counting march's frame bit for bit
    vis.set_counting(False)
    assert np.array_equal(vis.draw(), img)
    vis.close()


def test_bricks_nonpersistent_and_chains(gpu, oracle_lib):
    """Non-persistent sources sample the functor-applied field; persistent ones
    apply the chain per sample."""
    from oracle import oracle

    spec = scenes.prod_scene(n=48, fb=(320, 200), n_particles=5000)
    vols = [
        dataclasses.replace(spec.volumes[0], chain="mul(2.0)|add(-0.5)", persistent=False),
        dataclasses.replace(spec.volumes[1], chain="length|pow(0.5)"),
    ]
    spec = dataclasses.replace(spec, volumes=vols)
    vis, img, fimg = brick_frame(spec, gpu, F32)
    ora = oracle.render_post(spec, 32, F32)
    ora64 = oracle.render_post(spec, 64, oracle.BRICK_RAW)
    compare_post("chains", spec, img, fimg, vis, F32, ora, ora64)
    # the signature-compiled march with the chains interpreted per sample: the same frame
    vis.set_counting(False)
    assert np.array_equal(vis.draw(), img)
    vis.close()


def test_form_switch_needs_update(gpu):
    from isaac_amd import visualization as V
    from isaac_amd._lib import IsaacHipError

    spec = scenes.small_scene(n=32, fb=(64, 64))
    vis = V.from_scene(spec, gpu.index or 0)
    vis.do_visualization()
    vis.set_render_form(vis.FORM_BRICKS, F16)
    with pytest.raises(IsaacHipError):
        vis.draw()
    vis.update_buffers()
    a = vis.draw()
    vis.set_render_form(vis.FORM_COMBINED, F32)
    vis.update_buffers()
    b = vis.draw()
    assert a.shape == b.shape
    vis.close()


@pytest.mark.parametrize("dtype", [F16, BF16, F32])
def test_precision_workload_full_size(gpu, oracle_lib, dtype):
    """512^3 Gaussian mixture, 3840x2160, step 0.25: the GPU image against the FP64
    oracle on every 16th pixel of every 16th row (SSAO off, so pixels are independent)."""
    from oracle import oracle

    spec = _precision_spec()
    vis, img, fimg = brick_frame(spec, gpu, dtype)
    ora64 = _precision_fp64()
    sel = (slice(None, None, 16), slice(None, None, 16))
    a, b = fimg[sel], ora64.rgba_f[sel]
    e = np.abs(a.astype(np.float64) - b.astype(np.float64))
    p = oracle.psnr_float(a, b)
    d8 = np.abs(img[sel].astype(np.int32) - ora64.rgba[sel].astype(np.int32))
    print(f"precision512 {NAMES[dtype]}: PSNR vs FP64 {p:.2f} dB on {a.shape[0] * a.shape[1]} px, "
          f"max abs {e.max():.4g}, mean abs {e.mean():.3g}, RGBA8 within 2/255 on {(d8.max(axis=2) <= 2).mean() * 100:.3f}%",
          flush=True)
    assert p >= PSNR_FLOOR_DB
    vis.close()


_CACHE = {}


def _precision_spec():
    if "spec" not in _CACHE:
        _CACHE["spec"] = scenes.precision_scene()
    return _CACHE["spec"]


def _precision_fp64():
    from oracle import oracle

    if "fp64" not in _CACHE:
        _CACHE["fp64"] = oracle.render_post(_precision_spec(), 64, oracle.BRICK_RAW, stride=16)
    return _CACHE["fp64"]
