"""The learned occlusion pass on the GPU.

  * forward and gradients of the FP16-MFMA network against the float64 oracle
    (oracle/aonet_oracle.py); tolerances are those of FP16 activations
  * on-GPU training from frames rendered with the plain SSAO pass lowers the loss
  * the trained pass is gated on the final image: >= 40 dB PSNR against the
    plain pass on held-out poses, and well above the image without occlusion
"""
import dataclasses

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu


def _net_vis(gpu, spec=None):
    from isaac_amd import aonet
    from isaac_amd import visualization as V

    spec = spec or scenes.small_scene(n=32, fb=(64, 64))
    vis = V.from_scene(spec, gpu.index or 0)
    aonet.check_param_count()
    return vis


def _inputs(seed, n=2, h=32, w=48):
    from oracle import aonet_oracle as A

    rng = np.random.default_rng(seed)
    depth = rng.uniform(200, 600, size=(n, h, w))
    depth[:, :5, :9] = 3.4e38
    nrm = rng.standard_normal((n, h, w, 3))
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    return A.features(depth, nrm, 256.0), rng.uniform(0, 1, size=(n, h, w))


def test_param_layout_matches_oracle(gpu):
    from isaac_amd import aonet
    from oracle import aonet_oracle as A

    assert A.param_layout()[1] == aonet.param_count()
    assert np.array_equal(A.init_params(7), aonet.init_params(7))


def test_forward_matches_fp64_oracle(gpu):
    from isaac_amd import aonet
    from oracle import aonet_oracle as A

    vis = _net_vis(gpu)
    p = aonet.init_params(11)
    vis.aonet_set_params(p)
    x, _ = _inputs(1)
    # the library sees FP16 features: give the oracle the same rounded inputs
    x16 = x.astype(np.float16).astype(np.float64)
    got = vis.aonet_forward(x)
    want, _ = A.forward(p.astype(np.float64), x16)
    err = np.abs(got - want)
    scale = np.abs(want).max()
    print(f"forward: max |err| {err.max():.4g}, mean {err.mean():.3g}, output scale {scale:.3g}", flush=True)
    assert err.max() <= 2e-2 * (1.0 + scale)
    assert np.corrcoef(got.ravel(), want.ravel())[0, 1] > 0.9999
    vis.close()


def test_gradients_match_fp64_oracle(gpu):
    from isaac_amd import aonet
    from oracle import aonet_oracle as A

    vis = _net_vis(gpu)
    p = aonet.init_params(12)
    vis.aonet_set_params(p)
    x, t = _inputs(2)
    x16 = x.astype(np.float16).astype(np.float64)
    loss, g = vis.aonet_loss_grad(x, t)
    loss_ref, g_ref = A.loss_and_grad(p.astype(np.float64), x16, t.astype(np.float32).astype(np.float64))
    assert abs(loss - loss_ref) <= 1e-2 * loss_ref, (loss, loss_ref)
    cos = float(g @ g_ref / (np.linalg.norm(g) * np.linalg.norm(g_ref)))
    print(f"gradient: loss {loss:.6g} vs {loss_ref:.6g}, cosine {cos:.6f}", flush=True)
    assert cos > 0.999
    layout, _ = A.param_layout()
    for name, wo, ws, bo, bl in layout:
        n = int(np.prod(ws))
        for lo, hi, what in ((wo, wo + n, "w"), (bo, bo + bl, "b")):
            ref = g_ref[lo:hi]
            rel = np.linalg.norm(g[lo:hi] - ref) / max(np.linalg.norm(ref), 1e-12)
            assert rel < 0.03, (name, what, rel)
    vis.close()


def test_render_path_occlusion_matches_network(gpu):
    """The render path's learned occlusion map (pools and head fused into the
    convolutions, frame padded to multiples of 8 with background) equals
    clamp(network(features of its own G-buffer)) evaluated in float64."""
    from isaac_amd import aonet
    from oracle import aonet_oracle as A

    # close pose, opaque particles: about a third of the 100 x 60 frame is surface
    spec = scenes.prod_scene(n=48, fb=(100, 60), n_particles=50_000, pose=scenes.FILL_POSE)
    p0 = spec.particles[0]
    spec.particles[0] = dataclasses.replace(p0, tf_points=[(0, (1.0, 1.0, 1.0, 1.0)), (1024, (1.0, 0.8, 0.2, 1.0))],
                                            constant_radius=1.0)
    vis = _net_vis(gpu, spec)
    p = aonet.init_params(5)
    vis.aonet_set_params(p)
    vis.set_ao_mode(vis.AO_LEARNED)
    vis.update_buffers()
    vis.draw()
    _, depth, normal, ao = vis.gbuffer()
    scale = max(int(np.float32(g) * np.float32(s)) for g, s in zip(spec.global_size, spec.scale))
    H, W = depth.shape
    Hp, Wp = (H + 7) // 8 * 8, (W + 7) // 8 * 8
    d = np.full((1, Hp, Wp), np.finfo(np.float32).max, np.float64)
    nrm = np.zeros((1, Hp, Wp, 3))
    d[0, :H, :W] = depth
    nrm[0, :H, :W] = normal
    assert np.count_nonzero(depth < 3e38) > 0.2 * depth.size  # the frame has geometry
    x16 = A.features(d, nrm, float(scale)).astype(np.float16).astype(np.float64)
    want = np.clip(A.forward(p.astype(np.float64), x16)[0][0, :H, :W], 0.0, 1.0)
    err = np.abs(ao.astype(np.float64) - want)
    print(f"render-path occlusion vs float64 network: max |err| {err.max():.4g}, mean {err.mean():.3g}", flush=True)
    assert err.max() <= 2e-2
    vis.close()


def _evaluation_scene():
    # particles and an iso-surface give the depth structure the occlusion term depends on
    return scenes.prod_scene(n=128, fb=(1920, 1080), n_particles=250_000)


@pytest.fixture(scope="module")
def trained(gpu):
    """The network trained on 20,000 triples (1250 steps x 16 crops of 128^2) of
    16 production-style 128^3 scenes (fields, particles, clipping planes) whose seeds
    differ from the evaluation scene's, at random poses at 1920x1080, then set on a
    context of the evaluation scene for the gate."""
    from isaac_amd import aonet
    from isaac_amd import visualization as V

    ts = aonet.TrainingSet(gpu.index or 0, fb=(1920, 1080), n=128, particles=250_000)
    ts.vis.aonet_set_params(aonet.init_params(0))
    losses = aonet.train(ts.vis, steps=1250, seed=1, log_every=250, prepare=ts.prepare)
    params = ts.vis.aonet_get_params()
    ts.close()
    vis = V.from_scene(_evaluation_scene(), gpu.index or 0)
    vis.update_buffers()
    vis.aonet_set_params(params)
    yield vis, losses
    vis.close()


def test_training_lowers_loss(trained):
    vis, losses = trained
    first, last = float(np.mean(losses[:20])), float(np.mean(losses[-20:]))
    print(f"training mse: first 20 steps {first:.5f}, last 20 steps {last:.5f}", flush=True)
    assert last < 0.5 * first


def test_learned_pass_gate(trained):
    """Held-out poses (the training poses are random draws of another seed)."""
    from isaac_amd import aonet

    vis, _ = trained
    poses = [scenes.DEFAULT_POSE, scenes.FILL_POSE, scenes.random_pose(np.random.default_rng(99))]
    r = aonet.gate(vis, poses, floor_db=40.0, min_gain_db=3.0)
    print(f"learned occlusion: image PSNR {r['image_psnr']:.2f} dB (floor 40), occlusion map PSNR "
          f"{r['ao_psnr']:.2f} dB; without occlusion the image is at {r['image_psnr_no_ao']:.2f} dB; "
          f"gain over no occlusion (worst pose) {r['gain_db']:.2f} dB", flush=True)
    # what the network buys over the method without it, on a scene it was not trained on
    assert r["gain_db"] >= 3.0
    assert r["passed"], r
    assert vis.ao_mode == vis.AO_LEARNED


def test_gate_rejects_a_network_that_adds_nothing(gpu):
    """All-zero parameters give an occlusion term of 0 everywhere: the image equals the
    one without occlusion, which can sit above the 40 dB floor, but it gains 0 dB, so
    the gate falls back to the plain pass."""
    from isaac_amd import aonet
    from isaac_amd import visualization as V

    vis = V.from_scene(_evaluation_scene(), gpu.index or 0)
    vis.update_buffers()
    vis.aonet_set_params(np.zeros(aonet.param_count(), np.float32))
    poses = [scenes.DEFAULT_POSE, scenes.FILL_POSE]
    r = aonet.gate(vis, poses, floor_db=40.0, min_gain_db=3.0)
    print(f"zero network: image {r['image_psnr']:.2f} dB, without occlusion {r['image_psnr_no_ao']:.2f} dB, "
          f"gain {r['gain_db']:.3f} dB", flush=True)
    assert abs(r["gain_db"]) < 0.01
    assert not r["passed"]
    assert vis.ao_mode == vis.AO_PLAIN
    vis.close()


def test_learned_pass_frame_time(trained):
    """Frame time with the learned pass next to the plain pass (reported, not gated)."""
    vis, _ = trained
    vis.set_ao_mode(vis.AO_PLAIN)
    times = {}
    for mode in (vis.AO_PLAIN, vis.AO_LEARNED):
        vis.set_ao_mode(mode)
        for _ in range(3):
            vis.draw(out=False)
        ssao, kern = [], []
        for _ in range(10):
            vis.draw(out=False)
            t = vis.timings()
            ssao.append(t["ssao_us"])
            kern.append(t["kernel_us"])
        times[mode] = (np.median(ssao), np.median(kern))
    print(f"occlusion pass: plain {times[0][0]:.1f} us, learned {times[1][0]:.1f} us; "
          f"draw: plain {times[0][1]:.1f} us, learned {times[1][1]:.1f} us", flush=True)
    assert times[1][0] > 0
