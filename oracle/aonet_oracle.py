"""CPU ORACLE -- TEST INFRASTRUCTURE ONLY: the occlusion network in float64 numpy.

The library's learned occlusion pass replaces the reference's SSAO pass
(SSAOKernel, lib/isaac/isaac_ssao_kernel.hpp:161-182: 7x7 depth taps at a
10-pixel stride, count of closer taps / 42) with a small U-Net evaluated on the
G-buffer.  The network is this project's addition; this module restates its
forward pass, loss and gradients in float64 so the GPU's FP16-MFMA kernels can
be checked.  It is pinned in tests/test_aonet_oracle.py against torch's CPU
autograd (float64) and against finite differences.

Layout conventions shared with the library:
  activations   (N, H, W, C), channels last
  conv weights  (Cout, 9, Cin), tap t = ky * 3 + kx reads pixel (y + ky - 1, x + kx - 1),
                zero padding outside the image
  head          1x1 convolution 16 -> 1, linear
  parameters    one flat float32 vector: for each conv layer W then b, then the head
"""
from __future__ import annotations

import numpy as np

FEATURES = 8  # depth/scale, background flag, normal xyz, 3 zero channels
# (name, cin, cout): encoder e*, bottleneck, decoder d* (input = upsampled + skip)
LAYERS = [
    ("e0a", FEATURES, 16), ("e0b", 16, 16),
    ("e1a", 16, 32), ("e1b", 32, 32),
    ("e2a", 32, 64), ("e2b", 64, 64),
    ("e3a", 64, 64), ("e3b", 64, 64),
    ("d2", 64 + 64, 64),
    ("d1", 64 + 32, 32),
    ("d0", 32 + 16, 16),
]
HEAD_IN = 16
DEPTH_BG = 3.0e38  # depths at or above this are background (no surface)


def param_layout():
    """[(name, w_offset, w_shape, b_offset, b_len)] and the total length."""
    out, off = [], 0
    for name, cin, cout in LAYERS:
        wlen = cout * 9 * cin
        out.append((name, off, (cout, 9, cin), off + wlen, cout))
        off += wlen + cout
    out.append(("head", off, (1, HEAD_IN), off + HEAD_IN, 1))
    off += HEAD_IN + 1
    return out, off


def init_params(seed: int) -> np.ndarray:
    """He-normal conv weights from PCG64(seed), zero biases (float32 vector)."""
    layout, total = param_layout()
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.zeros(total, np.float32)
    for name, wo, ws, bo, bl in layout:
        fan_in = ws[1] * ws[2] if name != "head" else HEAD_IN
        p[wo:wo + int(np.prod(ws))] = (rng.standard_normal(int(np.prod(ws))) * np.sqrt(2.0 / fan_in)).astype(np.float32)
    return p


def unpack(p):
    layout, _ = param_layout()
    d = {}
    for name, wo, ws, bo, bl in layout:
        d[name] = (np.asarray(p[wo:wo + int(np.prod(ws))], np.float64).reshape(ws), np.asarray(p[bo:bo + bl], np.float64))
    return d


def features(depth, normal, scale):
    """(N, H, W) depth and (N, H, W, 3) normal -> (N, H, W, 8) network input."""
    depth = np.asarray(depth, np.float64)
    bg = depth >= DEPTH_BG
    f = np.zeros(depth.shape + (FEATURES,), np.float64)
    f[..., 0] = np.where(bg, 0.0, depth / scale)
    f[..., 1] = bg
    f[..., 2:5] = normal
    return f


def _cols(x):
    N, H, W, C = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    return np.stack([xp[:, ky:ky + H, kx:kx + W, :] for ky in range(3) for kx in range(3)], axis=3)  # N H W 9 C


def conv(x, w, b):
    return np.einsum("nhwtc,otc->nhwo", _cols(x), w, optimize=True) + b


def conv_backward(x, w, dz):
    """Gradients of z = conv(x, w, b) given dz: (dx, dw, db)."""
    N, H, W, C = x.shape
    dw = np.einsum("nhwo,nhwtc->otc", dz, _cols(x), optimize=True)
    db = dz.sum(axis=(0, 1, 2))
    dcols = np.einsum("nhwo,otc->nhwtc", dz, w, optimize=True)
    dxp = np.zeros((N, H + 2, W + 2, C))
    for t in range(9):
        ky, kx = divmod(t, 3)
        dxp[:, ky:ky + H, kx:kx + W, :] += dcols[:, :, :, t, :]
    return dxp[:, 1:H + 1, 1:W + 1, :], dw, db


def pool(x):
    N, H, W, C = x.shape
    return x.reshape(N, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4))


def pool_backward(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_backward(dy):
    N, H, W, C = dy.shape
    return dy.reshape(N, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


def forward(p, x):
    """Returns (ao_raw (N, H, W), cache).  H and W must be multiples of 8."""
    P = unpack(p)
    c = {"x": x}
    relu = lambda z: np.maximum(z, 0.0)

    def layer(name, inp):
        z = conv(inp, *P[name])
        c[name + ":in"] = inp
        c[name] = relu(z)
        return c[name]

    s0 = layer("e0b", layer("e0a", x))
    s1 = layer("e1b", layer("e1a", pool(s0)))
    s2 = layer("e2b", layer("e2a", pool(s1)))
    bt = layer("e3b", layer("e3a", pool(s2)))
    d2 = layer("d2", np.concatenate([up(bt), s2], axis=3))
    d1 = layer("d1", np.concatenate([up(d2), s1], axis=3))
    d0 = layer("d0", np.concatenate([up(d1), s0], axis=3))
    wh, bh = P["head"]
    out = np.einsum("nhwc,c->nhw", d0, wh[0]) + bh[0]
    return out, c


def loss_and_grad(p, x, target):
    """Sum-of-squares loss sum((ao_raw - target)^2) and its gradient wrt the flat parameters."""
    P = unpack(p)
    out, c = forward(p, x)
    err = out - target
    loss = float(np.sum(err * err))
    dout = 2.0 * err
    g = {}
    wh, bh = P["head"]
    g["head"] = (np.einsum("nhw,nhwc->c", dout, c["d0"])[None, :], np.array([dout.sum()]))
    dy = {"d0": np.einsum("nhw,c->nhwc", dout, wh[0])}

    def back(name):
        dz = dy[name] * (c[name] > 0)
        dx, dw, db = conv_backward(c[name + ":in"], P[name][0], dz)
        g[name] = (dw, db)
        return dx

    def add(name, v):
        dy[name] = dy.get(name, 0.0) + v

    dx = back("d0")
    add("d1", up_backward(dx[..., :32]))
    add("e0b", dx[..., 32:])
    dx = back("d1")
    add("d2", up_backward(dx[..., :64]))
    add("e1b", dx[..., 64:])
    dx = back("d2")
    add("e3b", up_backward(dx[..., :64]))
    add("e2b", dx[..., 64:])
    add("e3a", back("e3b"))
    add("e2b", pool_backward(back("e3a")))
    add("e2a", back("e2b"))
    add("e1b", pool_backward(back("e2a")))
    add("e1a", back("e1b"))
    add("e0b", pool_backward(back("e1a")))
    add("e0a", back("e0b"))
    back("e0a")
    layout, total = param_layout()
    flat = np.zeros(total, np.float64)
    for name, wo, ws, bo, bl in layout:
        flat[wo:wo + int(np.prod(ws))] = g[name][0].reshape(-1)
        flat[bo:bo + bl] = g[name][1]
    return loss, flat


def ssao(depth):
    """The reference SSAO pass on one (H, W) depth image (lib/isaac/isaac_ssao_kernel.hpp:161-182)."""
    H, W = depth.shape
    occ = np.zeros((H, W))
    for i in range(-3, 4):
        for j in range(-3, 4):
            ys = np.clip(np.arange(H) + j * 10, 0, H - 1)
            xs = np.clip(np.arange(W) + i * 10, 0, W - 1)
            occ += depth[ys][:, xs] < depth
    return np.clip(occ / 42.0, 0.0, 1.0)
