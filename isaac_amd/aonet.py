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
from typing import Dict, List, Optional, Sequence

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
) -> List[float]:
    """`steps` Adam steps on crops of frames rendered with the plain pass at random
    poses (or cycling through `poses`).  Needs ambient occlusion enabled on `vis`
    and network parameters set.  Returns the per-step mean squared error."""
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
            out["image_psnr"] = min(out["image_psnr"], psnr(learned, plain, 255.0))
            out["ao_psnr"] = min(out["ao_psnr"], psnr(ao_learned, ao_plain, 1.0))
            out["image_psnr_no_ao"] = min(out["image_psnr_no_ao"], psnr(none, plain, 255.0))
    finally:
        vis.ambient_occlusion["isEnabled"] = ao_on
        vis._push_state()
        vis.set_ao_mode(mode)
    return out


def gate(vis, poses: Sequence, floor_db: float = 40.0) -> Dict[str, float]:
    """Select the learned pass only if its final image stays within `floor_db` PSNR of
    the plain pass on every pose; otherwise fall back to the plain pass."""
    r = compare(vis, poses)
    r["passed"] = r["image_psnr"] >= floor_db
    vis.set_ao_mode(vis.AO_LEARNED if r["passed"] else vis.AO_PLAIN)
    return r


def check_param_count() -> None:
    n = int(_lib.lib().isaac_hip_aonet_param_count())
    if n != param_count():
        raise _lib.IsaacHipError(f"parameter layout mismatch: library {n}, python {param_count()}")
