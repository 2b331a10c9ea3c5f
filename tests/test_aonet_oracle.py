"""Pins of the occlusion-network oracle (oracle/aonet_oracle.py), on the CPU:
torch's float64 autograd for the forward pass and every parameter gradient,
central finite differences, and the SSAO restatement against the C++ oracle."""
import numpy as np
import pytest

from oracle import aonet_oracle as A


def _torch_forward(p, x):
    import torch
    import torch.nn.functional as F

    P = {}
    layout, _ = A.param_layout()
    pt = torch.tensor(np.asarray(p, np.float64), requires_grad=True)
    for name, wo, ws, bo, bl in layout:
        w = pt[wo:wo + int(np.prod(ws))].reshape(ws)
        b = pt[bo:bo + bl]
        if name != "head":
            # (Cout, 9, Cin) with t = ky*3 + kx  ->  (Cout, Cin, 3, 3)
            w = w.reshape(ws[0], 3, 3, ws[2]).permute(0, 3, 1, 2)
        P[name] = (w, b)
    xt = torch.tensor(np.asarray(x, np.float64)).permute(0, 3, 1, 2)

    def c(name, v):
        return F.relu(F.conv2d(v, P[name][0], P[name][1], padding=1))

    s0 = c("e0b", c("e0a", xt))
    s1 = c("e1b", c("e1a", F.avg_pool2d(s0, 2)))
    s2 = c("e2b", c("e2a", F.avg_pool2d(s1, 2)))
    bt = c("e3b", c("e3a", F.avg_pool2d(s2, 2)))
    up = lambda v: F.interpolate(v, scale_factor=2, mode="nearest")
    d2 = c("d2", torch.cat([up(bt), s2], 1))
    d1 = c("d1", torch.cat([up(d2), s1], 1))
    d0 = c("d0", torch.cat([up(d1), s0], 1))
    out = torch.einsum("nchw,c->nhw", d0, P["head"][0][0]) + P["head"][1][0]
    return out, pt


def _inputs(seed=0, n=2, h=16, w=24):
    rng = np.random.default_rng(seed)
    depth = rng.uniform(100, 400, size=(n, h, w))
    depth[:, :3, :5] = 3.4e38
    nrm = rng.standard_normal((n, h, w, 3))
    nrm /= np.linalg.norm(nrm, axis=-1, keepdims=True)
    x = A.features(depth, nrm, 256.0)
    t = rng.uniform(0, 1, size=(n, h, w))
    return x, t


def test_forward_matches_torch():
    import torch

    p = A.init_params(3)
    x, _ = _inputs()
    out, _ = A.forward(p, x)
    ref, _ = _torch_forward(p, x)
    assert np.allclose(out, ref.detach().numpy(), rtol=1e-10, atol=1e-10)


def test_gradients_match_torch_autograd():
    p = A.init_params(4)
    x, t = _inputs(1)
    loss, g = A.loss_and_grad(p, x, t)
    ref, pt = _torch_forward(p, x)
    import torch

    lt = ((ref - torch.tensor(t)) ** 2).sum()
    lt.backward()
    lv = float(lt.detach())
    assert abs(loss - lv) <= 1e-9 * abs(lv)
    gref = pt.grad.numpy()
    assert np.allclose(g, gref, rtol=1e-8, atol=1e-8 * np.abs(gref).max())


def test_gradients_finite_differences():
    p = A.init_params(5).astype(np.float64)
    x, t = _inputs(2, n=1, h=8, w=8)
    _, g = A.loss_and_grad(p, x, t)
    layout, total = A.param_layout()
    rng = np.random.default_rng(0)
    idx = list(rng.choice(total, 30, replace=False)) + [layout[-1][3], layout[0][3]]
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (A.loss_and_grad(pp, x, t)[0] - A.loss_and_grad(pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_features():
    depth = np.array([[[10.0, 3.4e38]]])
    nrm = np.array([[[[0.0, 0.6, 0.8], [0.0, 0.0, 0.0]]]])
    f = A.features(depth, nrm, 20.0)
    assert f.shape == (1, 1, 2, 8)
    assert list(f[0, 0, 0]) == [0.5, 0.0, 0.0, 0.6, 0.8, 0.0, 0.0, 0.0]
    assert list(f[0, 0, 1]) == [0.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]


def test_ssao_restatement_matches_cpp_oracle(oracle_lib):
    from isaac_amd import scenes

    spec = scenes.prod_scene(n=32, fb=(96, 64), n_particles=3000)
    res = oracle_lib.render(spec, 64)
    want = A.ssao(res.pre_depth)
    assert np.allclose(res.ao, want, atol=1e-6)
