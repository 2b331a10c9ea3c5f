"""GPU parity: the gfx950 render path against the CPU oracle on identical seeded inputs.

Bars:
  * combined RGBA8 textures (classification) ........ bit-exact
  * particle cell lists (counts, offsets, perm) ..... bit-exact
  * RGBA8 framebuffer ................................ |diff| <= 2/255 per channel on >= 99.9 % of pixels
  * executed sample / cell / sphere-test counts ...... equal (they define samples/s)
  * PSNR against the FP64 oracle is printed for every image.
"""
import dataclasses

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 2
FRACTION = 0.999


def _vis(spec, gpu):
    from isaac_amd import visualization as V

    return V.from_scene(spec, gpu.index or 0)


def gpu_frame(spec, gpu, counting=True):
    vis = _vis(spec, gpu)
    vis.set_counting(counting)
    img = vis.do_visualization()
    vis.synchronize()
    return vis, img


def compare(name, spec, gpu_img, ora, ora64=None):
    from oracle import oracle

    d = np.abs(gpu_img.astype(np.int32) - ora.rgba.astype(np.int32))
    ok = (d.max(axis=2) <= TOL).mean()
    msg = f"{name}: within {TOL}/255 on {ok * 100:.4f}% of pixels, max diff {d.max()}"
    if ora64 is not None:
        msg += f", PSNR gpu-vs-fp64 {oracle.psnr(gpu_img, ora64.rgba):.2f} dB, oracle32-vs-fp64 {oracle.psnr(ora.rgba, ora64.rgba):.2f} dB"
    print(msg, flush=True)
    assert ok >= FRACTION, msg
    return ok


def test_small_scene_parity(gpu, oracle_lib):
    from oracle import oracle

    spec = scenes.small_scene()
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32, want_textures=True)
    ora64 = oracle.render(spec, 64)
    vol, iso = vis.combined_textures()
    assert np.array_equal(vol, ora.vol_tex), "classification volume texture differs"
    assert np.array_equal(iso, ora.iso_tex), "classification iso texture differs"
    compare("small64", spec, img, ora, ora64)
    c = vis.counters()
    # the march weighs the eight texels with fmas where the reference nests lerps: a
    # ray's opacity may cross 0.99 one step apart in rare rays
    assert abs(c["vol_samples"] - ora.counters["vol_samples"]) <= 1e-4 * ora.counters["vol_samples"], (c, ora.counters)
    assert c["vol_rays"] == ora.counters["vol_rays"]
    assert (img[..., 3] > 0).sum() > 10000


def _reduced_prod(n=64, fb=(480, 270), particles=20000, visible_tf=False, pose=scenes.DEFAULT_POSE):
    spec = scenes.prod_scene(n=n, fb=fb, n_particles=particles, pose=pose)
    if visible_tf:
        p = spec.particles[0]
        spec.particles[0] = dataclasses.replace(
            p, tf_points=[(0, (1.0, 1.0, 1.0, 0.0)), (1024, (1.0, 0.8, 0.2, 1.0))], constant_radius=0.5
        )
    return spec


@pytest.mark.parametrize(
    "label,kw",
    [
        ("prod64-default", dict()),
        ("prod64-fill", dict(pose=scenes.FILL_POSE)),
        ("prod64-dense-particles", dict(visible_tf=True)),
    ],
)
def test_reduced_production_parity(gpu, oracle_lib, label, kw):
    from oracle import oracle

    spec = _reduced_prod(**kw)
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32, want_textures=True)
    ora64 = oracle.render(spec, 64)
    vol, iso = vis.combined_textures()
    assert np.array_equal(vol, ora.vol_tex)
    assert np.array_equal(iso, ora.iso_tex)
    color, depth, normal, ao = vis.gbuffer()
    compare(label, spec, img, ora, ora64)
    # depth after the solid passes: equal wherever both are finite surfaces
    both = (ora.depth < 3e38) & (depth < 3e38)
    assert (np.abs(depth[both] - ora.depth[both]) <= 1e-3 * np.maximum(1.0, np.abs(ora.depth[both]))).mean() >= FRACTION
    assert (np.abs(ao - ora.ao) <= 1e-6).mean() >= FRACTION
    c = vis.counters()
    for k in ("vol_samples", "iso_samples", "iso_hits", "cells_visited", "vol_rays", "iso_rays", "part_rays"):
        rel = abs(c[k] - ora.counters[k]) / max(1, ora.counters[k])
        assert rel <= 1e-3, (k, c[k], ora.counters[k])


def test_binning_bit_exact_full_size(gpu, oracle_lib):
    """2,000,000 uniform particles on a 64^3 cell grid of a 256^3 box."""
    from oracle import oracle

    spec = _reduced_prod(n=256, fb=(64, 64), particles=2_000_000)
    spec = dataclasses.replace(spec, volumes=[spec.volumes[0]])
    vis, _ = gpu_frame(spec, gpu, counting=False)
    counts, offs, perm = vis.cell_list(0)
    p = spec.particles[0]
    cell_of, ocounts, ooffs, operm, rel = oracle.bin_particles(p.pos, spec.local_size, spec.particle_grid)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(offs.astype(np.int64), ooffs)
    assert np.array_equal(perm.astype(np.int64), operm)
    # stable order: brute force
    ref = np.argsort(cell_of, kind="stable")
    assert np.array_equal(perm, ref)


def test_visible_records_and_full_cell_list(gpu, oracle_lib):
    """update_buffers builds only the visible particles' records; the full list built
    on request is bit-exact and renders the same frame as the visible build."""
    from oracle import oracle

    spec = _reduced_prod(n=64, fb=(320, 200), particles=50_000, pose=scenes.FILL_POSE)
    vis = _vis(spec, gpu)
    img = vis.do_visualization()
    visible = int((spec.particles[0].attr[:, 0] >= 0.995).sum())
    assert 0 < visible < 50_000
    compare("visible-records", spec, img, oracle.render(spec, 32))
    counts, offs, perm = vis.cell_list(0)
    p = spec.particles[0]
    cell_of, ocounts, ooffs, operm, rel = oracle.bin_particles(p.pos, spec.local_size, spec.particle_grid)
    assert np.array_equal(counts, ocounts)
    assert np.array_equal(offs.astype(np.int64), ooffs)
    assert np.array_equal(perm.astype(np.int64), operm)
    # the full build replaced the records: the frame is the same
    assert np.array_equal(vis.draw(), img)
    # a second request reuses the full list; a new update builds visible records again
    assert np.array_equal(vis.cell_list(0)[2], perm)
    vis.update_buffers()
    assert np.array_equal(vis.draw(), img)
    vis.close()


@pytest.mark.parametrize("mode", list(range(9)))
def test_render_modes(gpu, oracle_lib, mode):
    from oracle import oracle

    spec = dataclasses.replace(_reduced_prod(n=48, fb=(200, 150), particles=5000, visible_tf=True), render_mode=mode)
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32)
    compare(f"mode{mode}", spec, img, ora)


@pytest.mark.parametrize("mode", [2, 3, 4, 5])
def test_dither_modes(gpu, oracle_lib, mode):
    """Blue-noise dither masks 4^3, 8^3, 16^3 (and a mode past 4, which uses the
    16^3 mask): combined textures bit-exact, frame within the pixel bar."""
    from oracle import oracle

    spec = dataclasses.replace(_reduced_prod(n=40, fb=(160, 120), particles=2000), dither_mode=mode)
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32, want_textures=True)
    vol, iso = vis.combined_textures()
    assert np.array_equal(vol, ora.vol_tex), "classification volume texture differs"
    assert np.array_equal(iso, ora.iso_tex), "classification iso texture differs"
    compare(f"dither{mode}", spec, img, ora)


@pytest.mark.parametrize(
    "label,changes",
    [
        ("nearest", dict(interpolation=False)),
        ("dither0", dict(dither_mode=0)),
        ("viewport", dict(viewport=(37, 21, 150, 90))),
        ("two-clips", dict(clipping=[((0.1, 0.0, 0.0), (1.0, 0.2, 0.0)), ((0.0, 0.0, 0.3), (0.0, -0.3, 1.0))])),
        ("scaled", dict(scale=(1.0, 1.5, 0.75))),
        ("step", dict(step=1.3)),
        ("ao-weight", dict(ao_weight=0.8)),
    ],
)
def test_state_variants(gpu, oracle_lib, label, changes):
    from oracle import oracle

    spec = dataclasses.replace(_reduced_prod(n=48, fb=(200, 150), particles=5000, visible_tf=True), **changes)
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32)
    compare(label, spec, img, ora)


def test_functor_chains_and_nonpersistent(gpu, oracle_lib):
    from oracle import oracle

    spec = _reduced_prod(n=40, fb=(160, 120), particles=3000, visible_tf=True)
    v0, v1 = spec.volumes
    spec.volumes[0] = dataclasses.replace(v0, chain="mul(1.5)|add(-0.1)", persistent=False)
    spec.volumes[1] = dataclasses.replace(v1, chain="mul(2,1,0.5)|sum|pow(0.7)")
    spec.particles[0] = dataclasses.replace(spec.particles[0], chain="mul(0.9)|add(0.1)")
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32, want_textures=True)
    vol, iso = vis.combined_textures()
    # pow goes through the device's powf: allow a handful of LUT-index flips
    assert (vol != ora.vol_tex).any(axis=3).mean() < 1e-3
    compare("chains", spec, img, ora)


@pytest.mark.parametrize("pose", ["default", "fill"])
def test_production_full_size(gpu, oracle_lib, pose):
    """The benchmark workload itself (256^3 x 3 sources, 2M particles, 1920x1080)
    at both of its poses (the fill pose executes 272 M samples per frame)."""
    from oracle import oracle

    spec = scenes.prod_scene(pose=scenes.DEFAULT_POSE if pose == "default" else scenes.FILL_POSE)
    vis, img = gpu_frame(spec, gpu)
    img2 = vis.do_visualization()
    assert np.array_equal(img, img2), "two renders of the same frame differ"
    c = vis.counters()
    # the march over the binary16 copy of the volume texture renders the same frame and
    # executes the same steps
    vis.set_tuning(vis.TUNE_TEX_LAYOUT, 16)
    assert np.array_equal(vis.do_visualization(), img), "texture layout 16"
    assert vis.counters() == c
    ora = oracle.render(spec, 32)
    compare(f"prod256-{pose}", spec, img, ora)
    print("gpu counters", c, "oracle", ora.counters, flush=True)
    assert abs(c["vol_samples"] - ora.counters["vol_samples"]) <= 1e-3 * ora.counters["vol_samples"]


@pytest.mark.parametrize("form", ["combined", "bricks-f16"])
def test_schedules_do_not_change_the_frame(gpu, oracle_lib, form):
    """Every march schedule (tile order over the XCDs, occupancy target, prefetch,
    workgroups of 4 waves or of one wave, loads by 32-bit offset or 64-bit address)
    renders the same frame; the counting march executes the same steps in every
    tile order and workgroup size."""
    spec = _reduced_prod(n=64, fb=(330, 250), particles=20000, visible_tf=True)
    vis = _vis(spec, gpu)
    if form != "combined":
        vis.set_render_form(vis.FORM_BRICKS, vis.F16)
    vis.update_buffers()
    ref = ref_counts = None
    for threads in ((256, 64) if form == "combined" else (256,)):
        vis.set_tuning(vis.TUNE_MARCH_THREADS, threads)
        for tiles in (0, 2, 4, 8, 16):
            vis.set_tuning(vis.TUNE_MARCH_TILES, tiles)
            for waves, prefetch in ((0, 0), (6, 0), (8, 0), (0, 1)):
                vis.set_tuning(vis.TUNE_MARCH_WAVES, waves)
                vis.set_tuning(vis.TUNE_MARCH_PREFETCH, prefetch)
                for offsets in (32, 64):
                    vis.set_tuning(vis.TUNE_MARCH_OFFSETS, offsets)
                    img = vis.draw()
                    if ref is None:
                        ref = img
                    assert np.array_equal(img, ref), (threads, tiles, waves, prefetch, offsets)
                vis.set_tuning(vis.TUNE_MARCH_OFFSETS, 32)
            vis.set_counting(True)
            assert np.array_equal(vis.draw(), ref), (threads, tiles)
            c = {k: v for k, v in vis.counters().items() if k != "march_wave_steps"}
            vis.set_counting(False)
            if ref_counts is None:
                ref_counts = c
            assert c == ref_counts, (threads, tiles)
    vis.close()


@pytest.mark.parametrize("n,linear,pose", [(64, True, "default"), (64, True, "fill"), (37, True, "default"),
                                           (37, False, "fill")])
@pytest.mark.parametrize("layout", [8, 16])
def test_texture_layouts_render_the_same_frame(gpu, oracle_lib, n, linear, pose, layout):
    """The march reading the 8^3-brick copy of the combined texture (layout 8: apron
    texels and partial bricks at a side of 37 hold clamped texels) or the binary16
    copy of its volume RGBA (layout 16; the nearest sampler keeps reading the combined
    texture) renders the frame and G-buffer of the linear texture bit for bit, at
    every occupancy target, with prefetch (layout 16) and with the counting march,
    and executes the same steps; at a side of 37 the frame is also checked against
    the oracle."""
    from oracle import oracle

    p = scenes.FILL_POSE if pose == "fill" else scenes.DEFAULT_POSE
    spec = dataclasses.replace(_reduced_prod(n=n, fb=(330, 250), particles=20000, visible_tf=True, pose=p),
                               interpolation=linear)
    vis = _vis(spec, gpu)
    vis.update_buffers()
    vis.set_counting(True)
    ref = vis.draw()
    ref_g = [a.copy() for a in vis.gbuffer()]
    ref_c = {k: v for k, v in vis.counters().items()}
    vis.set_counting(False)
    vis.set_tuning(vis.TUNE_TEX_LAYOUT, layout)
    vis.update_buffers()
    for waves, prefetch in ((0, 0), (6, 0), (8, 0)) + (((0, 1),) if layout == 16 else ()):
        vis.set_tuning(vis.TUNE_MARCH_WAVES, waves)
        vis.set_tuning(vis.TUNE_MARCH_PREFETCH, prefetch)
        img = vis.draw()
        assert np.array_equal(img, ref), (waves, prefetch)
        for a, b in zip(vis.gbuffer(), ref_g):
            assert np.array_equal(a, b), (waves, prefetch)
    vis.set_tuning(vis.TUNE_MARCH_WAVES, 0)
    vis.set_tuning(vis.TUNE_MARCH_PREFETCH, 0)
    vis.set_counting(True)
    assert np.array_equal(vis.draw(), ref)
    assert vis.counters() == ref_c
    vis.close()
    if n == 37:
        compare(f"prod{n}-{pose} layout {layout}", spec, ref, oracle.render(spec, 32))


def test_texture_layout_change_needs_update(gpu):
    from isaac_amd._lib import IsaacHipError

    vis = _vis(_reduced_prod(n=16, fb=(32, 32), particles=100), gpu)
    vis.update_buffers()
    vis.draw()
    for bad in (4, 32):
        with pytest.raises(IsaacHipError):
            vis.set_tuning(vis.TUNE_TEX_LAYOUT, bad)
    vis.set_tuning(vis.TUNE_TEX_LAYOUT, 8)
    with pytest.raises(IsaacHipError):
        vis.draw()
    vis.update_buffers()
    vis.draw()
    vis.close()


@pytest.mark.parametrize("pose", ["default", "fill"])
def test_iso_hit_depth_matches_oracle(gpu, oracle_lib, pose):
    """Iso-surface hit depths, interpolated between the hit step and the step
    before it (lib/isaac/isaac_iso_kernel.hpp:84-92), against the oracle: the march
    forms the previous step's iso sum only when the hit needs it, which must give
    the same depths as forming it at every step (a wrong previous sum moves a hit
    by up to one step, about 1e-3 of the depth)."""
    from oracle import oracle

    p = scenes.FILL_POSE if pose == "fill" else scenes.DEFAULT_POSE
    spec = _reduced_prod(n=64, fb=(640, 480), particles=0, pose=p)
    # iso threshold 0.1: a surface over a few thousand pixels at both poses
    density = dataclasses.replace(spec.volumes[0], iso=0.1)
    spec = dataclasses.replace(spec, particles=[], volumes=[density], ao_enabled=False)
    vis, img = gpu_frame(spec, gpu)
    ora = oracle.render(spec, 32)
    gb = [a.copy() for a in vis.gbuffer()]
    depth = gb[1]
    # the march over the binary16 copy forms the iso sums from the combined texture at the
    # flagged steps only: the same hits, depths and normals
    vis.set_tuning(vis.TUNE_TEX_LAYOUT, 16)
    assert np.array_equal(vis.do_visualization(), img)
    for a, b in zip(vis.gbuffer(), gb):
        assert np.array_equal(a, b)
    vis.close()
    both = (ora.depth < 3e38) & (depth < 3e38)
    assert both.sum() > 3000, both.sum()
    rel = np.abs(depth[both] - ora.depth[both]) / np.abs(ora.depth[both])
    print(f"iso hits {both.sum()}: median rel depth diff {np.median(rel):.3g}, 99th pct {np.quantile(rel, 0.99):.3g}, "
          f"max {rel.max():.3g}", flush=True)
    assert (rel <= 1e-5).mean() >= 0.99
    assert np.array_equal(depth >= 3e38, ora.depth >= 3e38) or ((depth >= 3e38) != (ora.depth >= 3e38)).mean() < 1e-3


def test_tuning_rejects_unknown_values(gpu):
    from isaac_amd._lib import IsaacHipError

    vis = _vis(_reduced_prod(n=16, fb=(32, 32), particles=100), gpu)
    for key, bad in ((vis.TUNE_MARCH_TILES, 3), (vis.TUNE_MARCH_TILES, 32), (vis.TUNE_MARCH_WAVES, 5),
                     (vis.TUNE_MARCH_THREADS, 128), (vis.TUNE_MARCH_OFFSETS, 16), (99, 0)):
        with pytest.raises(IsaacHipError):
            vis.set_tuning(key, bad)
    for tiles in (0, 2, 4, 8, 16):
        vis.set_tuning(vis.TUNE_MARCH_TILES, tiles)
    vis.close()


@pytest.mark.parametrize("mode", [0, 1, 3])
def test_fused_shading_matches_separate_kernel(gpu, mode):
    """In render modes 0, 1 and 3 the occlusion kernel shades its own pixels: the frame,
    G-buffer and float output equal those of the separate shading kernel bit for bit."""
    spec = dataclasses.replace(_reduced_prod(n=48, fb=(333, 201), particles=5000, visible_tf=True), render_mode=mode)
    vis = _vis(spec, gpu)
    vis.update_buffers()
    out = {}
    for fused in (1, 0):
        vis.set_tuning(vis.TUNE_FUSED_SHADING, fused)
        img = vis.draw()
        out[fused] = (img, [a.copy() for a in vis.gbuffer()])
        if not fused:
            assert vis.timings()["shade_us"] > 0
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
    vis.close()


def test_timing_event_modes(gpu):
    """Timing events between all kernels (1), around the march only (2) or none (0):
    the same frame each time; get_timings reports what was recorded and zeros for the rest."""
    from isaac_amd._lib import IsaacHipError

    vis = _vis(_reduced_prod(n=32, fb=(96, 64), particles=2000), gpu)
    ref = vis.do_visualization()
    t = vis.timings()
    # (the occlusion kernel shades its pixels in this render mode: ssao_us holds the shading)
    assert t["march_us"] > 0 and t["ssao_us"] > 0 and t["optimization_buffer_us"] > 0
    for mode in (2, 0, 1):
        vis.set_tuning(vis.TUNE_TIMING_EVENTS, mode)
        assert np.array_equal(vis.do_visualization(), ref), mode
        t = vis.timings()
        if mode == 0:
            assert all(v == 0 for v in t.values()), t
        elif mode == 2:
            assert t["march_us"] > 0 and t["ssao_us"] == 0 and t["optimization_buffer_us"] == 0, t
        else:
            assert t["march_us"] > 0 and t["ssao_us"] > 0 and t["optimization_buffer_us"] > 0, t
    with pytest.raises(IsaacHipError):
        vis.set_tuning(vis.TUNE_TIMING_EVENTS, 3)
    vis.close()


@pytest.mark.parametrize("fb", [(1, 1), (17, 9), (300, 1), (1, 200), (129, 65)])
def test_odd_framebuffers(gpu, oracle_lib, fb):
    """Framebuffers of one pixel, one row or column, and sizes that leave partial
    tiles and super-tiles, in both render forms."""
    from oracle import oracle

    spec = _reduced_prod(n=32, fb=fb, particles=3000, visible_tf=True)
    for form in ("combined", "bricks"):
        vis = _vis(spec, gpu)
        if form == "bricks":
            vis.set_render_form(vis.FORM_BRICKS, vis.F32)
        vis.update_buffers()
        img = vis.draw()
        vis.close()
        assert img.shape == (fb[1], fb[0], 4)
        if form == "combined":
            compare(f"fb {fb}", spec, img, oracle.render(spec, 32))
        else:
            compare(f"fb {fb} bricks", spec, img, oracle.render_post(spec, 32, dtype=oracle.BRICK_F32))


def test_multi_source_weighting(gpu, oracle_lib):
    """Three volume sources with different weights, transfer functions and iso
    thresholds plus an invisible particle source: the classification divides by the
    volume sources' weights, the march multiplies by all sources' weights, the iso
    texel is the argmax of alpha / threshold / 2 (closed forms in
    tests/test_oracle_weighting.py).  Textures bit-exact, frame within the bar."""
    from oracle import oracle
    from test_oracle_weighting import ISO_CASES, ISO_TEXEL, VOL_TEXEL, _iso_scene, _weighting_scene

    for dither in (0, 1):
        spec = _weighting_scene(dither)
        vis, img = gpu_frame(spec, gpu)
        ora = oracle.render(spec, 32, want_textures=True)
        vol, iso = vis.combined_textures()
        assert np.array_equal(vol, ora.vol_tex)
        assert np.array_equal(iso, ora.iso_tex)
        if dither == 0:
            assert (vol == np.round(VOL_TEXEL * 255).astype(np.uint8)).all()
            assert (iso == ISO_TEXEL).all()
        compare(f"weights-dither{dither}", spec, img, ora)
        vis.close()
    for sources, expect in ISO_CASES:
        spec = _iso_scene(sources)
        vis, img = gpu_frame(spec, gpu)
        vol, iso = vis.combined_textures()
        assert (iso == np.array(expect, np.uint8)).all(), (sources, np.unique(iso.reshape(-1, 4), axis=0))
        vis.close()
