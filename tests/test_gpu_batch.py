"""The batched form against the per-context render path: every frame of a batch
(sub-volumes of uneven size x views) must equal, byte for byte, the frame the same
context renders on its own with the same camera -- for the eager persistent
launch, the hipGraph replay (replayed several times with other GPU work between)
and the one-launch-per-frame reference, with plain and with learned occlusion."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIDES = [16, 24, 32, 40, 28]
W, H = 96, 64
NVIEWS = 3


@pytest.fixture(scope="module")
def setup(gpu):
    from isaac_amd import batch as B

    vises, _ = B.make_subvolumes(gpu.index or 0, count=len(SIDES), seed=3, fb=(W, H), dtype=1, sides=SIDES)
    projs, mvs = B.views_for(vises, NVIEWS, seed=3)
    yield vises, projs, mvs
    for v in vises:
        v.close()


def _reference(vises, projs, mvs):
    out = []
    for f in range(len(vises) * NVIEWS):
        v = vises[f // NVIEWS]
        out.append(v.draw(projection=projs[f], modelview_=mvs[f]))
    return np.stack(out)


def _check_modes(vises, projs, mvs, label):
    from isaac_amd import batch as B

    ref = _reference(vises, projs, mvs)
    assert (ref[..., 3] > 0).mean() > 0.05, "frames should show something"
    b = B.Batch(vises, W, H)
    b.set_views(projs, mvs, bg=vises[0].background_color)
    for mode, name in ((B.PERSISTENT, "persistent"), (B.PER_FRAME, "per-frame"), (B.GRAPH, "graph")):
        b.render(mode)
        got = b.images()
        diff = np.abs(got.astype(int) - ref.astype(int)).max()
        print(f"{label} {name}: max diff {diff}", flush=True)
        assert np.array_equal(got, ref), (label, name)
    # graph replays, with the contexts' own rendering in between
    for k in range(3):
        _reference(vises[:1], projs[:NVIEWS], mvs[:NVIEWS])
        b.render(B.GRAPH)
        assert np.array_equal(b.images(), ref), (label, "replay", k)
    # other schedules give the same frames
    b.set_tuning(B.TUNE_CHUNK, 1)
    b.set_tuning(B.TUNE_GROUPS_PER_CU, 2)
    b.render(B.GRAPH)
    assert np.array_equal(b.images(), ref), (label, "chunk 1")
    b.close()


def test_batch_matches_per_context_plain_ssao(setup):
    vises, projs, mvs = setup
    for v in vises:
        v.set_ao_mode(v.AO_PLAIN)
    _check_modes(vises, projs, mvs, "plain")


def test_batch_matches_per_context_learned_occlusion(setup):
    from isaac_amd import aonet

    vises, projs, mvs = setup
    p = aonet.init_params(5)
    for v in vises:
        v.aonet_set_params(p)
        v.set_ao_mode(v.AO_LEARNED)
    try:
        _check_modes(vises, projs, mvs, "learned")
    finally:
        for v in vises:
            v.set_ao_mode(v.AO_PLAIN)


def test_batch_rejects_mixed_state(setup):
    from isaac_amd import batch as B
    from isaac_amd._lib import IsaacHipError

    vises, projs, mvs = setup
    vises[1].source_weight = [1.0, 2.0]
    vises[1].update_buffers()
    try:
        b = B.Batch(vises, W, H)
        with pytest.raises(IsaacHipError):
            b.set_views(projs, mvs)
        b.close()
    finally:
        vises[1].source_weight = [2.0, 2.0]
        vises[1].update_buffers()


def _particle_contexts(gpu, sides=(32, 24)):
    """Brick-form contexts with visible particles (same steering, different sizes)."""
    import dataclasses

    from isaac_amd import scenes
    from isaac_amd import visualization as V

    vises = []
    for n in sides:
        spec = scenes.prod_scene(n=n, fb=(W, H), n_particles=400)
        p = spec.particles[0]
        spec.particles[0] = dataclasses.replace(
            p, tf_points=[(0, (1.0, 1.0, 1.0, 0.0)), (1024, (1.0, 0.8, 0.2, 1.0))], constant_radius=0.5
        )
        spec = dataclasses.replace(spec, clipping=[], camera=scenes.FILL_POSE)
        v = V.from_scene(spec, gpu.index or 0)
        v.set_render_form(v.FORM_BRICKS, v.F16)
        v.update_buffers()
        vises.append(v)
    return vises


def test_batch_refuses_reallocated_contexts(gpu):
    """A context whose particle buffers grow (update_buffers reallocates them) or whose
    stream is replaced makes the next batch render fail with ERR_STATE instead of
    reading freed memory; after set_views the batch matches the per-context frames."""
    import torch

    from isaac_amd import batch as B
    from isaac_amd import visualization as V
    from isaac_amd._lib import IsaacHipError, lib

    vises = _particle_contexts(gpu)
    projs, mvs = B.views_for(vises, 2, seed=9)
    try:
        b = B.Batch(vises, W, H)
        b.set_views(projs, mvs)
        b.render(B.PERSISTENT)
        assert np.array_equal(b.images(), np.stack([vises[f // 2].draw(projection=projs[f], modelview_=mvs[f])
                                                    for f in range(4)]))
        # 4x the particles: the cell-list buffers of context 0 are reallocated
        dev = torch.device("cuda", gpu.index or 0)
        old = vises[0].particle_sources[0]
        n = old.pos.shape[0] * 4
        g = torch.Generator(device="cpu").manual_seed(4)
        pos = (torch.rand((n, 3), generator=g) * 31.0).to(dev)
        attr = torch.rand((n, 1), generator=g).to(dev)
        vises[0].particle_sources[0] = V.ParticleSource("p", pos, attr, None, 0.5)
        vises[0].update_buffers()
        torch.cuda.synchronize(dev)
        with pytest.raises(IsaacHipError, match="changed since"):
            b.render(B.PERSISTENT)
        with pytest.raises(IsaacHipError, match="changed since"):
            b.render(B.GRAPH)
        b.set_views(projs, mvs)
        for mode in (B.PERSISTENT, B.GRAPH):
            b.render(mode)
            ref = np.stack([vises[f // 2].draw(projection=projs[f], modelview_=mvs[f]) for f in range(4)])
            assert np.array_equal(b.images(), ref), mode
        # a new stream on context 0 (the batch keeps a stream of its own)
        assert lib().isaac_hip_set_stream(vises[0]._ctx, None) == 0
        with pytest.raises(IsaacHipError, match="changed since"):
            b.render(B.GRAPH)
        b.set_views(projs, mvs)
        b.render(B.GRAPH)
        assert np.array_equal(b.images(), ref)
        b.close()
    finally:
        for v in vises:
            v.close()


def test_batches_of_different_sizes_share_contexts(setup):
    """Two batches over the same contexts at different frame sizes, rendered
    alternately from their captured graphs with learned occlusion: each keeps its own
    occlusion workspace, so every replay equals its per-frame render."""
    from isaac_amd import aonet
    from isaac_amd import batch as B

    vises, projs, mvs = setup
    p = aonet.init_params(7)
    for v in vises:
        v.aonet_set_params(p)
        v.set_ao_mode(v.AO_LEARNED)
    try:
        batches, refs = [], []
        for (w, h) in ((W, H), (64, 40)):
            b = B.Batch(vises, w, h)
            b.set_views(projs, mvs)
            b.render(B.PER_FRAME)
            refs.append(b.images())
            batches.append(b)
        for k in range(3):
            for b, ref in zip(batches, refs):
                b.render(B.GRAPH)
                assert np.array_equal(b.images(), ref), (b.width, b.height, k)
        for b in batches:
            b.close()
    finally:
        for v in vises:
            v.set_ao_mode(v.AO_PLAIN)
