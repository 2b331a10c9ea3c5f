"""isaac_amd — MI355X (gfx950) native in-situ ray-casting renderer.

The render path (classification, particles, iso-surface, SSAO, shading,
volume) runs in hand-written HIP kernels in isaac_amd/libisaac_hip.so behind
the C-ABI include/isaac_hip.h.  `isaac_amd.visualization.IsaacVisualization`
mirrors the reference's per-rank IsaacVisualization interface on top of it.
"""
__version__ = "0.1.0"

from . import scenes  # noqa: F401
