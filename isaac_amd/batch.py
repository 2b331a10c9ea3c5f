"""Batched form: many sub-volumes x many camera views in one persistent launch
(include/isaac_hip.h, "batched form"), with an autotuner for its launch
schedule, and the generator of the batched workload (48 sub-volumes of uneven
size, 8 views each at 960x540)."""
from __future__ import annotations

import ctypes
import itertools
import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import IsaacHipError, lib

PERSISTENT, GRAPH, PER_FRAME = 0, 1, 2
TUNE_CHUNK, TUNE_GROUPS_PER_CU, TUNE_AO_CHUNK = 0, 1, 2


class Batch:
    """Frames = visualizations x views, rendered together.  The visualizations
    (IsaacVisualization, brick form, buffers updated) share context 0's steering."""

    def __init__(self, vises: Sequence, width: int, height: int):
        self.vises = list(vises)
        self.width, self.height = int(width), int(height)
        arr = (ctypes.c_void_p * len(self.vises))(*[v._ctx.value for v in self.vises])
        out = ctypes.c_void_p()
        rc = lib().isaac_hip_batch_create(arr, len(self.vises), self.width, self.height, ctypes.byref(out))
        if rc < 0:
            raise IsaacHipError(f"isaac_hip_batch_create failed: {_lib.ERRORS.get(rc, rc)}")
        self._b = out
        self.nviews = 0

    def close(self) -> None:
        if getattr(self, "_b", None):
            lib().isaac_hip_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str) -> None:
        if rc < 0:
            m = lib().isaac_hip_batch_last_error(self._b)
            raise IsaacHipError(f"{what} failed: {_lib.ERRORS.get(rc, rc)} {m.decode() if m else ''}".strip())

    def set_tuning(self, key: int, value: int) -> None:
        self._check(lib().isaac_hip_batch_set_tuning(self._b, int(key), int(value)), "batch_set_tuning")

    def set_views(self, projections: np.ndarray, modelviews: np.ndarray, bg=(0.0, 0.0, 0.0, 0.0)) -> None:
        """projections / modelviews: (n_vis * n_views, 16) column-major, frame-major."""
        p = np.ascontiguousarray(projections, np.float64).reshape(-1, 16)
        m = np.ascontiguousarray(modelviews, np.float64).reshape(-1, 16)
        if p.shape != m.shape or p.shape[0] % len(self.vises):
            raise ValueError("one projection and one modelview per frame (visualizations x views)")
        self.nviews = p.shape[0] // len(self.vises)
        b = np.asarray(bg, np.float32)
        self._check(lib().isaac_hip_batch_set_views(self._b, self.nviews, p.ctypes.data, m.ctypes.data, b.ctypes.data),
                    "batch_set_views")

    def render(self, mode: int = GRAPH) -> None:
        self._check(lib().isaac_hip_batch_render(self._b, int(mode)), "batch_render")

    def synchronize(self) -> None:
        self._check(lib().isaac_hip_batch_synchronize(self._b), "batch_synchronize")

    def last_time_us(self) -> float:
        us = ctypes.c_double()
        self._check(lib().isaac_hip_batch_last_time_us(self._b, ctypes.byref(us)), "batch_last_time_us")
        return float(us.value)

    def images(self) -> np.ndarray:
        """Host copy of all frames: (nframes, H, W, 4) uint8."""
        ptr, n = ctypes.c_void_p(), ctypes.c_int32()
        self._check(lib().isaac_hip_batch_get_images(self._b, ctypes.byref(ptr), ctypes.byref(n)), "batch_get_images")
        self.synchronize()
        out = np.zeros((n.value, self.height, self.width, 4), np.uint8)
        v0 = self.vises[0]
        v0._d2h(out, ptr.value)
        return out

    def images_device(self):
        """Device pointer and frame count of the images (no copy)."""
        ptr, n = ctypes.c_void_p(), ctypes.c_int32()
        self._check(lib().isaac_hip_batch_get_images(self._b, ctypes.byref(ptr), ctypes.byref(n)), "batch_get_images")
        return ptr.value, n.value


# ------------------------------------------------------------------ autotuner
_TUNE_CACHE: Dict[Tuple, Dict[str, int]] = {}


def autotune(batch: Batch, repeats: int = 3, candidates: Optional[Dict[str, Sequence[int]]] = None) -> Dict:
    """Time the graph-replayed batch under each (chunk, groups per CU) schedule and
    keep the fastest; cached per (frames, size, sub-volume sizes)."""
    cand = candidates or {"chunk": (1, 2, 4, 8), "groups_per_cu": (2, 4, 8)}
    key = (batch.width, batch.height, batch.nviews, tuple(tuple(v.local_size) for v in batch.vises))
    if key in _TUNE_CACHE:
        c = _TUNE_CACHE[key]
        batch.set_tuning(TUNE_CHUNK, c["chunk"])
        batch.set_tuning(TUNE_GROUPS_PER_CU, c["groups_per_cu"])
        return {"choice": dict(c), "cached": True, "us": {}}
    times = {}
    for chunk, gpc in itertools.product(cand["chunk"], cand["groups_per_cu"]):
        batch.set_tuning(TUNE_CHUNK, chunk)
        batch.set_tuning(TUNE_GROUPS_PER_CU, gpc)
        batch.render(GRAPH)  # capture + warm
        batch.synchronize()
        ts = []
        for _ in range(repeats):
            batch.render(GRAPH)
            ts.append(batch.last_time_us())
        times[(chunk, gpc)] = float(np.median(ts))
    best = min(times, key=times.get)
    choice = {"chunk": best[0], "groups_per_cu": best[1]}
    _TUNE_CACHE[key] = choice
    batch.set_tuning(TUNE_CHUNK, choice["chunk"])
    batch.set_tuning(TUNE_GROUPS_PER_CU, choice["groups_per_cu"])
    return {"choice": choice, "cached": False, "us": {f"chunk={c},groups_per_cu={g}": t for (c, g), t in times.items()}}


# ------------------------------------------------------------------ batched workload
SIDES = (32, 64, 96, 128, 192, 256)


def subvolume_sides(count: int = 48, seed: int = 5) -> List[int]:
    rng = np.random.Generator(np.random.PCG64(seed))
    return [int(s) for s in rng.choice(SIDES, size=count)]


def _gaussians_gpu(n: int, rng: np.random.Generator, device, count: int = 8):
    import torch

    s = n / 64.0
    c = rng.uniform(0.0, n, size=(count, 3))
    sig = rng.uniform(4.0 * s, 12.0 * s, size=count)
    amp = rng.uniform(0.2, 1.0, size=count)
    ax = (torch.arange(n, device=device, dtype=torch.float32) + 0.5)
    out = torch.zeros((n, n, n), device=device, dtype=torch.float32)
    for k in range(count):
        g = [torch.exp(-((ax - float(c[k, d])) ** 2) / float(2.0 * sig[k] ** 2)) for d in range(3)]
        out += float(amp[k]) * g[2][:, None, None] * g[1][None, :, None] * g[0][None, None, :]
    return out


def _value_noise_gpu(n: int, rng: np.random.Generator, device, freq: int = 4, octaves: int = 2):
    """Sum of octaves of trilinearly interpolated lattice noise in [-1, 1]."""
    import torch
    import torch.nn.functional as Fn

    out = torch.zeros((n, n, n), device=device, dtype=torch.float32)
    amp = 1.0
    for o in range(octaves):
        f = freq * (2 ** o)
        lat = torch.from_numpy(rng.uniform(-1.0, 1.0, size=(1, 1, f + 1, f + 1, f + 1)).astype(np.float32)).to(device)
        out += amp * Fn.interpolate(lat, size=(n, n, n), mode="trilinear", align_corners=True)[0, 0]
        amp *= 0.5
    return out


def make_subvolumes(device_index: int = 0, count: int = 48, seed: int = 5, fb=(960, 540), dtype: int = 1,
                    sides: Optional[Sequence[int]] = None):
    """`count` IsaacVisualizations (brick form, `dtype` bricks), each one sub-volume of side
    U{32..256} (or `sides`): Gaussian-mixture density (iso 0.5) + float3 noise shown through
    `length`, SSAO on, trilinear, step 0.5.  Returns (vises, sides)."""
    import torch

    from . import scenes
    from . import visualization as V

    device = torch.device("cuda", device_index)
    rng = np.random.Generator(np.random.PCG64(seed))
    sides = list(sides) if sides is not None else subvolume_sides(count, seed)
    vises = []
    for i, n in enumerate(sides):
        dens = _gaussians_gpu(n, rng, device)
        dens = dens / torch.clamp(dens.max(), min=1e-6)
        vec = torch.stack([_value_noise_gpu(n, rng, device) for _ in range(3)], dim=-1).contiguous()
        vols = [V.VolumeSource("density", dens.contiguous()), V.VolumeSource("vector", vec)]
        vis = V.IsaacVisualization(f"sub{i}", fb, (n, n, n), (n, n, n), (max(1, n // 4),) * 3, (0, 0, 0), vols, [], [],
                                   device=device_index)
        vis.functions = ["idem", "length"]
        vis.update_functions()
        vis.set_transfer_points(0, scenes.DENSITY_TF)
        vis.set_transfer_points(1, scenes.VECTOR_TF)
        vis.update_transfer()
        vis.iso_threshold = [0.5, 0.0]
        vis.source_weight = [2.0, 2.0]
        vis.interpolation = True
        vis.step = 0.5
        vis.ambient_occlusion = {"isEnabled": True, "weight": 0.5}
        vis.set_render_form(vis.FORM_BRICKS, dtype)
        vis.update_buffers()
        vises.append(vis)
    torch.cuda.synchronize(device)
    return vises, sides


def poses_for(count: int, nviews: int = 8, seed: int = 5):
    """The seeded camera poses of `count` sub-volumes x `nviews` views, frame-major."""
    from . import scenes

    rng = np.random.Generator(np.random.PCG64(seed + 1000))
    return [scenes.random_pose(rng) for _ in range(count * nviews)]


def views_for(vises: Sequence, nviews: int = 8, seed: int = 5):
    """Seeded random poses per visualization; (projections, modelviews) frame-major."""
    from .visualization import modelview

    poses = poses_for(len(vises), nviews, seed)
    projs, mvs = [], []
    for f, p in enumerate(poses):
        projs.append(vises[f // nviews].projection)
        mvs.append(modelview(p.look_at, p.rotation, p.distance))
    return np.stack(projs), np.stack(mvs)
