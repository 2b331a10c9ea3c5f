"""Learned occlusion pass: parameter layout, initialisation and the on-GPU
training loop around the library's U-Net (include/isaac_hip.h, "learned
occlusion").

The network stands in for the reference's SSAO pass
(lib/isaac/isaac_ssao_kernel.hpp:161-182).  Training data are generated on the
fly: each step renders the scene at a fresh random pose with the plain pass,
cuts square crops out of the G-buffer (depth, normal -> plain occlusion term)
and takes one Adam step on them -- all on the GPU.  `gate` compares the final
image rendered with the learned pass against the plain pass on held-out poses
and keeps the learned pass only above a PSNR floor.
"""
from __future__ import annotations

import math
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _lib

FEATURES = 8
# (name, cin, cout) in parameter order
LAYERS = [
    ("e0a", FEATURES, 16), ("e0b", 16, 16),
    ("e1a", 16, 32), ("e1b", 32, 32),
    ("e2a", 32, 64), ("e2b", 64, 64),
    ("e3a", 64, 64), ("e3b", 64, 64),
    ("d2", 128, 64), ("d1", 96, 32), ("d0", 48, 16),
]
HEAD_IN = 16


# resolution level (divisor 2^level) of each layer
LEVELS = [0, 0, 1, 1, 2, 2, 3, 3, 2, 1, 0]


def param_count() -> int:
    return sum(co * 9 * ci + co for _, ci, co in LAYERS) + HEAD_IN + 1


def flops_per_frame(width: int, height: int) -> float:
    """Multiply-add FLOPs (2 per MAC) of one inference over a width x height frame
    (padded to multiples of 8, as the library runs it): the 3x3 layers and the head."""
    wp, hp = (width + 7) // 8 * 8, (height + 7) // 8 * 8
    macs = sum((wp >> lv) * (hp >> lv) * co * 9 * ci for (_, ci, co), lv in zip(LAYERS, LEVELS))
    return 2.0 * (macs + wp * hp * HEAD_IN)


def init_params(seed: int) -> np.ndarray:
    """He-normal convolution and head weights, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _, ci, co in LAYERS:
        out.append(rng.standard_normal(co * 9 * ci) * math.sqrt(2.0 / (9 * ci)))
        out.append(np.zeros(co))
    out.append(rng.standard_normal(HEAD_IN) * math.sqrt(2.0 / HEAD_IN))
    out.append(np.zeros(1))
    return np.concatenate(out).astype(np.float32)


def _set_pose(vis, pose) -> None:
    vis.look_at = np.asarray(pose.look_at, np.float64)
    vis.rotation = np.asarray(pose.rotation, np.float64)
    vis.distance = float(pose.distance)
    vis._update_modelview()


def train(
    vis,
    steps: int,
    poses: Optional[Sequence] = None,
    seed: int = 0,
    crops: int = 16,
    side: int = 128,
    lr: float = 2e-3,
    lr_final: float = 2e-4,
    log_every: int = 0,
    prepare: Optional[Callable[[int, np.random.Generator], None]] = None,
) -> List[float]:
    """`steps` Adam steps on crops of frames rendered with the plain pass at random
    poses (or cycling through `poses`).  Needs ambient occlusion enabled on `vis`
    and network parameters set.  `prepare(k, rng)`, if given, changes the scene
    before step k's frame (TrainingSet.prepare swaps fields, particles and the
    clipping plane); update_buffers runs after it.  Returns the per-step mean
    squared error."""
    from .scenes import random_pose

    rng = np.random.default_rng(seed)
    W, H = vis.framebuffer_size
    if side > W or side > H:
        raise ValueError("crop side exceeds the framebuffer")
    mode = vis.ao_mode
    vis.set_ao_mode(vis.AO_PLAIN)
    vis.update_buffers()
    losses = []
    try:
        for k in range(steps):
            pose = poses[k % len(poses)] if poses else random_pose(rng)
            _set_pose(vis, pose)
            if prepare is not None:
                prepare(k, rng)
                vis.update_buffers()
            vis.draw(out=False)
            xy = np.stack([rng.integers(0, W - side + 1, crops), rng.integers(0, H - side + 1, crops)], axis=1)
            vis.aonet_clear_batch()
            vis.aonet_add_crops(xy.astype(np.int32), side)
            # cosine decay from lr to lr_final
            t = k / max(1, steps - 1)
            cur = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * t))
            losses.append(vis.aonet_train_step(cur))
            if log_every and (k % log_every == 0 or k == steps - 1):
                print(f"occlusion net step {k}: mse {losses[-1]:.5f} lr {cur:.2e}", flush=True)
    finally:
        vis.set_ao_mode(mode)
    return losses


class TrainingSet:
    """A training context and `count` seeded scenes in the production workload's
    style, each a density field (sum of 8 Gaussians, sigma U[4,12] x n/64, amplitude
    U[0.2,1], normalised to 1, blended with w x lattice value noise, w ~ U[0,0.4],
    clamped to [0,1]; iso-surface at 0.5) and `particles` uniform particles of
    radius 0.5 with the particle transfer function.  prepare(k, rng) copies scene
    k % count into the context's sources and draws a clipping plane (none in 3 of
    10 steps, else a random normal through a point of [-0.5,0.5]^3): train(ts.vis,
    ..., prepare=ts.prepare) renders depth / normal / occlusion triples at random
    poses.  The seeds differ from those of the scenes the network is evaluated on."""

    def __init__(self, device_index: int = 0, fb=(1920, 1080), n: int = 128, count: int = 16,
                 particles: int = 250_000, seed: int = 11):
        import torch

        from . import batch as B
        from . import scenes
        from . import visualization as V

        device = torch.device("cuda", device_index)
        rng = np.random.Generator(np.random.PCG64(seed))
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
        self.fields, self.positions = [], []
        for _ in range(count):
            g = B._gaussians_gpu(n, rng, device)
            g = g / torch.clamp(g.max(), min=1e-6)
            v = B._value_noise_gpu(n, rng, device)
            v = (v - v.min()) / torch.clamp(v.max() - v.min(), min=1e-6)
            w = float(rng.uniform(0.0, 0.4))
            self.fields.append(torch.clamp((1.0 - 0.5 * w) * g + w * v, 0.0, 1.0).contiguous())
            pos = torch.rand((particles, 3), generator=gen, device=device) * float(n)
            self.positions.append(torch.clamp(pos, max=float(np.nextafter(np.float32(n), np.float32(0)))).contiguous())
        attr = torch.rand((particles, 1), generator=gen, device=device).contiguous()
        vol = V.VolumeSource("density", self.fields[0].clone())
        part = V.ParticleSource("particles", self.positions[0].clone(), attr, None, 0.5)
        vis = V.IsaacVisualization("occlusion-training", fb, (n, n, n), (n, n, n), (max(1, n // 4),) * 3, (0, 0, 0),
                                   [vol], [], [part], device=device_index)
        vis.functions = ["idem", "idem"]
        vis.update_functions()
        vis.set_transfer_points(0, scenes.DENSITY_TF)
        vis.set_transfer_points(1, scenes.PARTICLE_TF)
        vis.update_transfer()
        vis.iso_threshold = [0.5, 0.0]
        vis.source_weight = [2.0, 2.0]
        vis.interpolation = True
        vis.step = 0.5
        vis.ambient_occlusion = {"isEnabled": True, "weight": 0.5}
        vis.update_buffers()
        torch.cuda.synchronize(device)
        self.vis, self._vol, self._part = vis, vol, part

    def prepare(self, k: int, rng: np.random.Generator) -> None:
        import torch

        i = k % len(self.fields)
        self._vol.tensor.copy_(self.fields[i])
        self._part.pos.copy_(self.positions[i])
        torch.cuda.current_stream(self._vol.tensor.device).synchronize()
        vis = self.vis
        if rng.random() < 0.3:
            while vis.clipping:
                vis.remove_clipping(0)
        else:
            nrm = rng.normal(size=3)
            nrm /= np.linalg.norm(nrm)
            pos = rng.uniform(-0.5, 0.5, size=3)
            if vis.clipping:
                vis.edit_clipping(0, *pos, *nrm)
            else:
                vis.add_clipping(*pos, *nrm)

    def close(self) -> None:
        self.vis.close()


def psnr(a: np.ndarray, b: np.ndarray, peak: float) -> float:
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0.0 else 10.0 * math.log10(peak * peak / mse)


def compare(vis, poses: Sequence) -> Dict[str, float]:
    """Final RGBA8 images and occlusion maps with the learned pass against the plain
    pass on `poses`: worst-pose PSNR of each, and of the image without occlusion."""
    out = {"image_psnr": float("inf"), "ao_psnr": float("inf"), "image_psnr_no_ao": float("inf")}
    mode = vis.ao_mode
    ao_on = vis.ambient_occlusion["isEnabled"]
    try:
        for p in poses:
            _set_pose(vis, p)
            vis.set_ao_mode(vis.AO_PLAIN)
            plain = vis.draw()
            ao_plain = vis.gbuffer()[3]
            vis.set_ao_mode(vis.AO_LEARNED)
            learned = vis.draw()
            ao_learned = vis.gbuffer()[3]
            vis.ambient_occlusion["isEnabled"] = False
            vis._push_state()
            none = vis.draw()
            vis.ambient_occlusion["isEnabled"] = ao_on
            vis._push_state()
            p_learned, p_none = psnr(learned, plain, 255.0), psnr(none, plain, 255.0)
            out["image_psnr"] = min(out["image_psnr"], p_learned)
            out["ao_psnr"] = min(out["ao_psnr"], psnr(ao_learned, ao_plain, 1.0))
            out["image_psnr_no_ao"] = min(out["image_psnr_no_ao"], p_none)
            # what the learned pass adds over leaving the occlusion out, on the same pose
            out["gain_db"] = min(out.get("gain_db", float("inf")), p_learned - p_none)
    finally:
        vis.ambient_occlusion["isEnabled"] = ao_on
        vis._push_state()
        vis.set_ao_mode(mode)
    return out


def gate(vis, poses: Sequence, floor_db: float = 40.0, min_gain_db: float = 3.0) -> Dict[str, float]:
    """Select the learned pass only if, on every pose, its final image stays within
    `floor_db` PSNR of the plain pass AND is at least `min_gain_db` closer to it than
    the image without any occlusion term (a network that adds nothing gains 0 dB and
    is rejected); otherwise fall back to the plain pass."""
    r = compare(vis, poses)
    r["floor_db"], r["min_gain_db"] = floor_db, min_gain_db
    r["passed"] = r["image_psnr"] >= floor_db and r["gain_db"] >= min_gain_db
    vis.set_ao_mode(vis.AO_LEARNED if r["passed"] else vis.AO_PLAIN)
    return r


def check_param_count() -> None:
    n = int(_lib.lib().isaac_hip_aonet_param_count())
    if n != param_count():
        raise _lib.IsaacHipError(f"parameter layout mismatch: library {n}, python {param_count()}")
