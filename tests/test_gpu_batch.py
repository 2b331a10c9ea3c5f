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
