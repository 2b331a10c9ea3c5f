"""CPU ORACLE — TEST INFRASTRUCTURE ONLY (sort-last compositing).

Restates, in numpy, the two steps of the reference's IceT compositing that one
GPU's share of it consists of:

  composite_order   the distance sort feeding icetCompositeOrder
                    lib/isaac.hpp:2021-2048 (point, glm length, std::multimap)
  over_ubyte        IceT's ICET_COMPOSITE_MODE_BLEND on RGBA_UBYTE
                    (mode and format set at lib/isaac.hpp:821-823)
  composite         ordered blending associated as a pairwise tree over the
                    composite order (binary swap, lib/isaac.hpp:811, for
                    power-of-two counts)

IceT itself is not in the reference (INSTALL.md:120-141 clones it from its git
master, unpinned); its byte "over" is restated from its published source:
dest = (IceTUByte)(dest * (255 - src.alpha) / 255 + src), integer division.
Parity of over_ubyte and of the association against IceT is UNPINNED (no
reference test or golden image covers compositing); the tests pin them by what
the algebra fixes instead (opaque and empty operands, exact cases, the tree
against a level-by-level binary-swap simulation).  glm's mat4 * vec4 and
length association ((a + b) + (c + d)) follow glm's published source.

Only tests/ may import this module; the product never does.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np


def composite_order(positions: Sequence[Sequence[int]], local_sizes: Sequence[Sequence[int]],
                    global_size: Sequence[int], scale: Sequence[float], modelview: np.ndarray) -> List[int]:
    """lib/isaac.hpp:2021-2048.  modelview: 16 column-major doubles."""
    f32 = np.float32
    gs = [int(f32(global_size[i]) * f32(scale[i])) for i in range(3)]
    mx = max(gs)
    m = np.asarray(modelview, np.float64).reshape(16)
    dist = []
    for pos, loc in zip(positions, local_sizes):
        p = []
        for i in range(3):
            ps = int(f32(pos[i]) * f32(scale[i]))
            ls = int(f32(loc[i]) * f32(scale[i]))
            p.append((float(ps) + (float(ls) - float(gs[i])) / 2.0) / float(mx // 2))
        p.append(1.0)
        v = [(m[r] * p[0] + m[4 + r] * p[1]) + (m[8 + r] * p[2] + m[12 + r] * p[3]) for r in range(4)]
        dist.append(np.float32(np.sqrt((v[0] * v[0] + v[1] * v[1]) + (v[2] * v[2] + v[3] * v[3]))))
    # std::multimap<float, rank>: ascending, equal keys in insertion (rank) order
    return sorted(range(len(dist)), key=lambda r: (dist[r], r))


def over_ubyte(front: np.ndarray, back: np.ndarray) -> np.ndarray:
    """front over back on (..., 4) uint8 RGBA arrays."""
    f = front.astype(np.uint32)
    b = back.astype(np.uint32)
    fac = 255 - f[..., 3:4]
    return ((b * fac) // 255 + f).astype(np.uint8)  # the uint8 cast keeps the value mod 256


def composite(images: Sequence[np.ndarray], order: Sequence[int]) -> np.ndarray:
    """images[order[0]] over images[order[1]] over ..., pairwise tree: adjacent
    pairs of the current level are blended (front = lower position), an odd last
    one passes to the next level unchanged."""
    level = [np.asarray(images[o], np.uint8) for o in order]
    while len(level) > 1:
        nxt = [over_ubyte(level[i], level[i + 1]) for i in range(0, len(level) - 1, 2)]
        if len(level) % 2:
            nxt.append(level[-1])
        level = nxt
    return level[0].copy()
