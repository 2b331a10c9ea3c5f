"""Launch-schedule autotuner.

Times the march kernel of a visualization under each candidate schedule
(include/isaac_hip.h, "tuning": march occupancy target, next-step prefetch,
tile order over the XCDs and, in the combined form, the texture layout: linear,
8^3 bricks or, where the sources stay put between frames, the binary16 copy, and the
workgroup size: 4 waves or one) on the caller's own frames and keeps the fastest.  Results do not
depend on the schedule, only speed does; choices are cached per workload key.
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
    "tex_layout": (0, 8, 16),
    "march_threads": (256, 64),
}


def _key(vis) -> Tuple:
    return (vis.framebuffer_size, vis.local_size, len(vis.volume_sources), len(vis.particle_sources),
            bool(vis.interpolation), vis.render_form)


def _apply(vis, choice: Dict[str, int]) -> None:
    vis.set_tuning(vis.TUNE_MARCH_WAVES, choice["march_waves"])
    vis.set_tuning(vis.TUNE_MARCH_PREFETCH, choice["march_prefetch"])
    vis.set_tuning(vis.TUNE_MARCH_TILES, choice.get("march_tiles", 4))
    vis.set_tuning(vis.TUNE_MARCH_THREADS, choice.get("march_threads", 256))
    layout = choice.get("tex_layout", 0)
    if layout != vis.tex_layout:
        vis.set_tuning(vis.TUNE_TEX_LAYOUT, layout)
        vis.update_buffers()


def tune_march(vis, poses: Sequence, frames: int = 4, use_cache: bool = True, update_each_frame: bool = True) -> Dict:
    """Pick the fastest march schedule over `poses` (frames per pose per candidate).
    With `update_each_frame` (the render loop updates the sources every frame) each
    timed frame runs update_buffers first and the score adds its classification time,
    which holds the brick layout's copy and the binary16 copy's stores.  Returns {"choice": {...}, "march_us":
    {candidate: score in us}}."""
    key = _key(vis)
    if use_cache and key in _CACHE:
        _apply(vis, _CACHE[key])
        return {"choice": dict(_CACHE[key]), "march_us": {}, "cached": True}
    results = {}
    combined = vis.render_form == vis.FORM_COMBINED
    layouts = CANDIDATES["tex_layout"] if combined else (vis.tex_layout,)
    if update_each_frame:
        # layout 16's binary16 copy costs the classification about what it saves the march
        # (measured at the 256^3 workload), so it is a candidate where the sources stay put
        layouts = tuple(x for x in layouts if x != 16) or (0,)
    threads = CANDIDATES["march_threads"] if combined else (256,)
    for layout, nt, waves, pf, tiles in itertools.product(layouts, threads, CANDIDATES["march_waves"],
                                                          CANDIDATES["march_prefetch"], CANDIDATES["march_tiles"]):
        if pf and (waves or layout == 8):
            continue  # the occupancy-targeted and brick-layout kernels have no prefetch variant
        choice = {"march_waves": waves, "march_prefetch": pf, "march_tiles": tiles, "tex_layout": layout,
                  "march_threads": nt}
        _apply(vis, choice)
        per_pose = []
        for p in poses:
            vis.look_at = np.asarray(p.look_at, np.float64)
            vis.rotation = np.asarray(p.rotation, np.float64)
            vis.distance = float(p.distance)
            vis._update_modelview()
            vis.draw(out=False)  # warm
            times = []
            for _ in range(frames):
                if update_each_frame:
                    vis.update_buffers()
                vis.draw(out=False)
                t = vis.timings()
                times.append(t["march_us"] + (t["optimization_buffer_us"] if update_each_frame else 0.0))
            per_pose.append(float(np.median(times)))
        # the mean over the poses of each pose's median: every pose weighs the same
        results[(layout, nt, waves, pf, tiles)] = (float(np.mean(per_pose)), per_pose)
    best = min(results, key=lambda k: results[k][0])
    choice = {"march_waves": best[2], "march_prefetch": best[3], "march_tiles": best[4], "tex_layout": best[0],
              "march_threads": best[1]}
    _CACHE[key] = choice
    _apply(vis, choice)

    def name(k):
        return f"layout={k[0]},threads={k[1]},waves={k[2]},prefetch={k[3]},tiles={k[4]}"

    return {"choice": choice, "cached": False,
            "march_us": {name(k): t[0] for k, t in results.items()},
            "march_us_per_pose": {name(k): t[1] for k, t in results.items()}}
