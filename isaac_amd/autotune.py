"""Launch-schedule autotuner.

Times the march kernel of a visualization under each candidate schedule
(include/isaac_hip.h, "tuning": march occupancy target, next-step prefetch,
tile order over the XCDs) on the caller's own frames and keeps the fastest.  Results do not depend on the
schedule, only speed does; choices are cached per workload key.
"""
from __future__ import annotations

import itertools
from typing import Dict, Sequence, Tuple

import numpy as np

_CACHE: Dict[Tuple, Dict[str, int]] = {}

CANDIDATES = {
    "march_waves": (0, 6, 8),
    "march_prefetch": (0, 1),
    "march_tiles": (0, 4, 8),
}


def _key(vis) -> Tuple:
    return (vis.framebuffer_size, vis.local_size, len(vis.volume_sources), len(vis.particle_sources),
            bool(vis.interpolation), vis.render_form if hasattr(vis, "render_form") else 0)


def _apply(vis, choice: Dict[str, int]) -> None:
    vis.set_tuning(vis.TUNE_MARCH_WAVES, choice["march_waves"])
    vis.set_tuning(vis.TUNE_MARCH_PREFETCH, choice["march_prefetch"])
    vis.set_tuning(vis.TUNE_MARCH_TILES, choice.get("march_tiles", 4))


def tune_march(vis, poses: Sequence, frames: int = 4, use_cache: bool = True) -> Dict:
    """Pick the fastest march schedule over `poses` (frames per pose per candidate).
    Returns {"choice": {...}, "march_us": {candidate: median us}}."""
    key = _key(vis)
    if use_cache and key in _CACHE:
        _apply(vis, _CACHE[key])
        return {"choice": dict(_CACHE[key]), "march_us": {}, "cached": True}
    results = {}
    for waves, pf, tiles in itertools.product(CANDIDATES["march_waves"], CANDIDATES["march_prefetch"],
                                              CANDIDATES["march_tiles"]):
        if pf and waves:
            continue  # the occupancy-targeted kernels have no prefetch variant
        choice = {"march_waves": waves, "march_prefetch": pf, "march_tiles": tiles}
        _apply(vis, choice)
        times = []
        for p in poses:
            vis.look_at = np.asarray(p.look_at, np.float64)
            vis.rotation = np.asarray(p.rotation, np.float64)
            vis.distance = float(p.distance)
            vis._update_modelview()
            vis.draw(out=False)  # warm
            for _ in range(frames):
                vis.draw(out=False)
                times.append(vis.timings()["march_us"])
        results[(waves, pf, tiles)] = float(np.median(times))
    best = min(results, key=results.get)
    choice = {"march_waves": best[0], "march_prefetch": best[1], "march_tiles": best[2]}
    _CACHE[key] = choice
    _apply(vis, choice)
    return {"choice": choice, "cached": False,
            "march_us": {f"waves={w},prefetch={p},tiles={s}": t for (w, p, s), t in results.items()}}
