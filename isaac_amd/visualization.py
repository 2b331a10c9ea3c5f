"""Python mirror of the reference's per-rank IsaacVisualization interface.

`IsaacVisualization` keeps the reference's constructor arguments, steering
state and setters (lib/isaac.hpp:565-1295) and its per-frame call
`do_visualization` (lib/isaac.hpp:1298-2104) for one rank: the buffer update
(source refresh, particle cell lists, classification) followed by the draw
callback (lib/isaac.hpp:2140-2543).  Everything runs on the gfx950 library
through the C-ABI in include/isaac_hip.h; PyTorch only provides device memory
for the sources.  There is no server, JSON, MPI or IceT layer.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from ._lib import IsaacHipError, check, lib

ISAAC_MAX_FUNCTORS = 3


# --------------------------------------------------------------------------- sources
class VolumeSource:
    """A volume source: `feature_dim`, `guard_size`, `persistent`, `get_name()`,
    `update(enabled, pointer)` (example/example.cpp:39-143) plus `tensor`, a
    CUDA float32 tensor of shape (Z+2g, Y+2g, X+2g[, C]) standing in for the
    reference's device-side `operator[]`."""

    feature_dim = 1
    guard_size = 0
    persistent = True

    def __init__(self, name: str, tensor, persistent: bool = True, guard_size: int = 0):
        import torch

        if not tensor.is_cuda or tensor.dtype != torch.float32:
            raise TypeError("volume source tensors must be CUDA float32")
        self.name = name
        self.tensor = tensor.contiguous()
        self.persistent = persistent
        self.guard_size = guard_size
        self.feature_dim = 1 if tensor.dim() == 3 else int(tensor.shape[3])

    def get_name(self) -> str:
        return self.name

    def update(self, enabled: bool, pointer=None) -> None:
        pass

    def descriptor(self) -> _lib.VolumeDesc:
        t = self.tensor
        g = self.guard_size
        Zg, Yg, Xg = int(t.shape[0]), int(t.shape[1]), int(t.shape[2])
        d = _lib.VolumeDesc()
        off = (g + g * Xg + g * Xg * Yg) * self.feature_dim * 4
        d.data = t.data_ptr() + off
        d.dtype = 0
        d.feature_dim = self.feature_dim
        d.stride[0], d.stride[1], d.stride[2] = 1, Xg, Xg * Yg
        d.guard_size = g
        d.persistent = 1 if self.persistent else 0
        return d


class ParticleSource:
    """A particle source: raw particles (local voxel coordinates) whose cell
    lists the library builds (the reference's getIterator(cell),
    example/example.cpp:148-228)."""

    def __init__(self, name: str, pos, attr, radius=None, constant_radius: float = 0.5):
        self.name = name
        self.pos = pos.contiguous()
        self.attr = attr.contiguous()
        self.radius = radius.contiguous() if radius is not None else None
        self.constant_radius = constant_radius
        self.feature_dim = int(attr.shape[1])

    def get_name(self) -> str:
        return self.name

    def update(self, enabled: bool, pointer=None) -> None:
        pass

    def descriptor(self) -> _lib.ParticleDesc:
        d = _lib.ParticleDesc()
        d.position = self.pos.data_ptr()
        d.attribute = self.attr.data_ptr()
        d.feature_dim = self.feature_dim
        d.radius = self.radius.data_ptr() if self.radius is not None else None
        d.constant_radius = self.constant_radius
        d.count = int(self.pos.shape[0])
        return d


# --------------------------------------------------------------------------- host helpers
def default_projection(w: int, h: int) -> np.ndarray:
    out = np.zeros(16, np.float64)
    lib().isaac_hip_default_projection(int(w), int(h), out.ctypes.data)
    return out


def modelview(look_at, rotation: np.ndarray, distance: float) -> np.ndarray:
    """updateModelview; `rotation` is the math 3x3 matrix R (rows = output axes)."""
    la = np.asarray(look_at, np.float64)
    rot = np.ascontiguousarray(np.asarray(rotation, np.float64).T).reshape(9)  # column-major
    out = np.zeros(16, np.float64)
    lib().isaac_hip_modelview(la.ctypes.data, rot.ctypes.data, float(distance), out.ctypes.data)
    return out


def parse_functor_chain(chain: str) -> Tuple[List[int], List[List[float]], int]:
    bc = np.zeros(3, np.int32)
    pr = np.zeros(12, np.float32)
    err = lib().isaac_hip_parse_functor_chain(chain.encode(), bc.ctypes.data, pr.ctypes.data)
    return bc.tolist(), pr.reshape(3, 4).tolist(), int(err)


def build_transfer_lut(points: Sequence[Tuple[int, Sequence[float]]], T: int) -> np.ndarray:
    pos = np.array([int(p) for p, _ in points], np.uint32)
    rgba = np.array([list(c) for _, c in points], np.float32).reshape(-1)
    lut = np.zeros((T, 4), np.float32)
    check(lib().isaac_hip_build_transfer_lut(T, len(points), pos.ctypes.data, rgba.ctypes.data, lut.ctypes.data))
    return lut


# --------------------------------------------------------------------------- visualization
class IsaacVisualization:
    """One rank's render state (IsaacVisualization<...>, lib/isaac.hpp:58-3036)."""

    def __init__(
        self,
        name: str,
        framebuffer_size: Tuple[int, int],
        global_size: Tuple[int, int, int],
        local_size: Tuple[int, int, int],
        local_particle_size: Tuple[int, int, int],
        position: Tuple[int, int, int],
        volume_sources: Sequence[VolumeSource],
        field_sources: Sequence[VolumeSource],
        particle_sources: Sequence[ParticleSource],
        scale: Tuple[float, float, float] = (1.0, 1.0, 1.0),
        transfer_size: int = 1024,
        device: int = 0,
        master: int = 0,
        server_url: Optional[str] = None,
        server_port: Optional[int] = None,
    ):
        self.name = name
        self.framebuffer_size = tuple(int(v) for v in framebuffer_size)
        self.global_size = tuple(int(v) for v in global_size)
        self.local_size = tuple(int(v) for v in local_size)
        self.local_particle_size = tuple(int(v) for v in local_particle_size)
        self.position = tuple(int(v) for v in position)
        self.scale = tuple(float(v) for v in scale)
        self.transfer_size = int(transfer_size)
        self.volume_sources = list(volume_sources)
        self.field_sources = list(field_sources)
        self.particle_sources = list(particle_sources)
        self.device = device
        n = len(self.volume_sources) + len(self.field_sources) + len(self.particle_sources)
        self.combined_source_list_size = n
        cfg = _lib.Config()
        cfg.framebuffer_size[:] = self.framebuffer_size
        cfg.global_size[:] = self.global_size
        cfg.local_size[:] = self.local_size
        cfg.local_particle_size[:] = self.local_particle_size
        cfg.position[:] = self.position
        cfg.scale[:] = self.scale
        cfg.transfer_size = self.transfer_size
        cfg.n_volume_sources = len(self.volume_sources)
        cfg.n_field_sources = len(self.field_sources)
        cfg.n_particle_sources = len(self.particle_sources)
        ctx = ctypes.c_void_p()
        check(lib().isaac_hip_create(ctypes.byref(cfg), device, ctypes.byref(ctx)), None, "isaac_hip_create")
        self._ctx = ctx
        # steering state with the reference's defaults (Appendix of defaults: lib/isaac.hpp:608-609,720-794,2987-2988)
        nv, nf = len(self.volume_sources), len(self.field_sources)
        self.functions = ["idem"] * nv + ["length"] * nf + ["idem"] * len(self.particle_sources)
        self.function_errors = [0] * n
        self.source_weight = [2.0] * n
        self.iso_threshold = [0.0] * n
        self.transfer_points: List[Dict[int, Tuple[float, float, float, float]]] = []
        for i in range(n):
            pos = np.zeros(2, np.uint32)
            rgba = np.zeros(8, np.float32)
            lib().isaac_hip_default_transfer(i, n, 1 if i >= nv + nf else 0, self.transfer_size, pos.ctypes.data, rgba.ctypes.data)
            self.transfer_points.append({int(pos[0]): tuple(rgba[:4].tolist()), int(pos[1]): tuple(rgba[4:].tolist())})
        self.interpolation = False
        self.step = 0.5
        self.dither_mode = 1
        self.render_mode = 0
        self.ambient_occlusion = {"isEnabled": False, "weight": 0.5}
        # field-source advection (lib/isaac.hpp:610,639-642)
        self.seed_points = 1000
        self.advection_step = 1.0
        self.advection_weight = 0.95
        self.advection_seeding_period = 1
        self.advection_seeding_duration = 1
        self.background_color = [0.0, 0.0, 0.0, 1.0]
        self.look_at = np.zeros(3)
        self.rotation = np.eye(3)
        self.distance = -4.5
        self.projection = default_projection(*self.framebuffer_size)
        self.modelview = modelview(self.look_at, self.rotation, self.distance)
        self.clipping: List[Tuple[Tuple[float, float, float], Tuple[float, float, float]]] = []
        self.viewport: Optional[Tuple[int, int, int, int]] = None  # IceT readback viewport (x, y, w, h)
        self.kernel_time = self.copy_time = self.buffer_time = self.optimization_buffer_time = 0.0
        self._sources_bound = False
        self.update_functions()

    # ---------------------------------------------------------------- lifetime
    def close(self) -> None:
        if getattr(self, "_ctx", None):
            lib().isaac_hip_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, communicator_setting=None) -> int:
        """No server connection in this library; always succeeds."""
        return 0

    # ---------------------------------------------------------------- setters
    def update_functions(self) -> None:
        for i, f in enumerate(self.functions):
            err = ctypes.c_int32(0)
            check(lib().isaac_hip_set_function(self._ctx, i, f.encode(), ctypes.byref(err)), self._ctx, "set_function")
            self.function_errors[i] = int(err.value)

    def update_transfer(self) -> None:
        for i, pts in enumerate(self.transfer_points):
            items = sorted(pts.items())
            pos = np.array([k for k, _ in items], np.uint32)
            rgba = np.array([v for _, v in items], np.float32).reshape(-1)
            check(
                lib().isaac_hip_set_transfer_points(self._ctx, i, len(items), pos.ctypes.data, rgba.ctypes.data),
                self._ctx,
                "set_transfer_points",
            )

    def set_transfer_points(self, idx: int, points: Sequence[Tuple[int, Sequence[float]]]) -> None:
        d: Dict[int, Tuple[float, ...]] = {}
        for k, v in points:
            d.setdefault(int(k), tuple(float(x) for x in v))  # std::map::insert keeps the first
        self.transfer_points[idx] = d

    def add_clipping(self, px, py, pz, nx, ny, nz) -> bool:
        ok = check(lib().isaac_hip_add_clipping(self._ctx, px, py, pz, nx, ny, nz), self._ctx, "add_clipping") == 1
        if ok:
            self.clipping.append(((px, py, pz), (nx, ny, nz)))
        return ok

    def edit_clipping(self, nr, px, py, pz, nx, ny, nz) -> bool:
        ok = check(lib().isaac_hip_edit_clipping(self._ctx, nr, px, py, pz, nx, ny, nz), self._ctx, "edit_clipping") == 1
        if ok and nr < len(self.clipping):
            self.clipping[nr] = ((px, py, pz), (nx, ny, nz))
        return ok

    def remove_clipping(self, nr: int) -> None:
        check(lib().isaac_hip_remove_clipping(self._ctx, nr), self._ctx, "remove_clipping")
        if nr < len(self.clipping):
            del self.clipping[nr]

    def update_position(self, position) -> None:
        self.position = tuple(int(v) for v in position)
        arr = (ctypes.c_int32 * 3)(*self.position)
        check(lib().isaac_hip_update_position(self._ctx, arr), self._ctx, "update_position")

    def update_local_size(self, local_size) -> None:
        self.local_size = tuple(int(v) for v in local_size)
        arr = (ctypes.c_int64 * 3)(*self.local_size)
        check(lib().isaac_hip_update_local_size(self._ctx, arr), self._ctx, "update_local_size")

    def update_local_particle_size(self, size) -> None:
        self.local_particle_size = tuple(int(v) for v in size)
        arr = (ctypes.c_int64 * 3)(*self.local_particle_size)
        check(lib().isaac_hip_update_local_particle_size(self._ctx, arr), self._ctx, "update_local_particle_size")

    # camera steering (the "rotation/position/distance" JSON keys, lib/isaac.hpp:1440-1513)
    def set_rotation_absolute(self, rotation: np.ndarray) -> None:
        self.rotation = np.asarray(rotation, np.float64)
        self._update_modelview()

    def rotate_relative(self, relative: np.ndarray) -> None:
        self.rotation = np.asarray(relative, np.float64) @ self.rotation
        self._update_modelview()

    def rotate_axis(self, axis, degrees: float) -> None:
        from .scenes import rotation_axis

        self.rotation = rotation_axis(axis, degrees) @ self.rotation
        self._update_modelview()

    def set_position_absolute(self, look_at) -> None:
        self.look_at = np.asarray(look_at, np.float64)
        self._update_modelview()

    def move_relative(self, translation) -> None:
        self.look_at = self.look_at + self.rotation.T @ np.asarray(translation, np.float64)
        self._update_modelview()

    def set_distance(self, distance: float) -> None:
        self.distance = float(distance)
        self._update_modelview()

    def distance_relative(self, delta: float) -> None:
        self.distance += float(delta)
        self._update_modelview()

    def _update_modelview(self) -> None:
        self.modelview = modelview(self.look_at, self.rotation, self.distance)

    def _push_state(self) -> None:
        """Hand the steering members to the library; skipped when none changed since
        the last push (these setters are the only writers of that state)."""
        ao = self.ambient_occlusion
        key = (tuple(float(w) for w in self.source_weight), tuple(float(t) for t in self.iso_threshold), float(self.step),
               bool(self.interpolation), int(self.dither_mode), bool(ao["isEnabled"]), float(ao["weight"]),
               int(self.render_mode))
        if self.field_sources:
            key += (float(self.advection_step), float(self.advection_weight), int(self.advection_seeding_period),
                    int(self.advection_seeding_duration), int(self.seed_points))
        if key == getattr(self, "_pushed", None):
            return
        c = self._ctx
        L = lib()
        for i in range(self.combined_source_list_size):
            check(L.isaac_hip_set_weight(c, i, float(self.source_weight[i])), c, "set_weight")
            check(L.isaac_hip_set_iso_threshold(c, i, float(self.iso_threshold[i])), c, "set_iso")
        check(L.isaac_hip_set_step(c, float(self.step)), c, "set_step")
        check(L.isaac_hip_set_interpolation(c, int(bool(self.interpolation))), c, "set_interpolation")
        check(L.isaac_hip_set_dither_mode(c, int(self.dither_mode)), c, "set_dither_mode")
        ao = self.ambient_occlusion
        check(L.isaac_hip_set_ao(c, int(bool(ao["isEnabled"])), float(ao["weight"])), c, "set_ao")
        check(L.isaac_hip_set_render_mode(c, int(self.render_mode)), c, "set_render_mode")
        if self.field_sources:
            check(L.isaac_hip_set_advection(c, float(self.advection_step), float(self.advection_weight),
                                            int(self.advection_seeding_period), int(self.advection_seeding_duration)),
                  c, "set_advection")
            check(L.isaac_hip_set_seed_points(c, int(self.seed_points)), c, "set_seed_points")
        self._pushed = key

    def _bind_sources(self, pointer=None) -> None:
        c = self._ctx
        nv = len(self.volume_sources) + len(self.field_sources)
        for i, s in enumerate(self.volume_sources + self.field_sources):
            enabled = self.source_weight[i] != 0.0 or self.iso_threshold[i] != 0.0
            s.update(enabled, pointer)
            d = s.descriptor()
            check(lib().isaac_hip_set_volume_source(c, i, ctypes.byref(d)), c, "set_volume_source")
        for j, s in enumerate(self.particle_sources):
            s.update(self.source_weight[nv + j] != 0.0, pointer)
            d = s.descriptor()
            check(lib().isaac_hip_set_particle_source(c, j, ctypes.byref(d)), c, "set_particle_source")

    # ---------------------------------------------------------------- per frame
    def update_buffers(self, pointer=None) -> None:
        """doVisualization's buffer section (lib/isaac.hpp:1805-1948)."""
        self._push_state()
        self._bind_sources(pointer)
        check(lib().isaac_hip_update_buffers(self._ctx), self._ctx, "update_buffers")

    def draw(self, out=None, viewport=None, projection=None, modelview_=None):
        """drawCallBack (lib/isaac.hpp:2140-2543).  `out`: None -> host numpy
        RGBA8 (H, W, 4); a CUDA uint8 tensor -> device copy; a host numpy array or
        (pinned) CPU tensor of H*W*4 bytes -> host copy into it; False -> no copy."""
        self._push_state()
        W, H = self.framebuffer_size
        proj = np.ascontiguousarray(projection if projection is not None else self.projection, np.float64)
        mv = np.ascontiguousarray(modelview_ if modelview_ is not None else self.modelview, np.float64)
        bg = np.asarray(self.background_color, np.float32)
        if viewport is None:
            viewport = self.viewport if self.viewport is not None else (0, 0, W, H)
        vp = np.asarray(viewport, np.int32)
        result = None
        if out is None:
            result = np.zeros((H, W, 4), np.uint8)
            ptr, dev = result.ctypes.data, 0
        elif out is False:
            ptr, dev = None, 0
        elif isinstance(out, np.ndarray):
            if out.nbytes != W * H * 4 or not out.flags["C_CONTIGUOUS"]:
                raise ValueError("host framebuffer: a contiguous array of width*height*4 bytes")
            ptr, dev = out.ctypes.data, 0
        elif not out.is_cuda:
            if out.numel() * out.element_size() != W * H * 4 or not out.is_contiguous():
                raise ValueError("host framebuffer: a contiguous tensor of width*height*4 bytes")
            ptr, dev = out.data_ptr(), 0
        else:
            ptr, dev = out.data_ptr(), 1
        check(
            lib().isaac_hip_render(self._ctx, proj.ctypes.data, mv.ctypes.data, bg.ctypes.data, vp.ctypes.data, ptr, dev),
            self._ctx,
            "render",
        )
        return result if out is None else out

    def do_visualization(self, meta_targets=None, pointer=None, redraw: bool = True, out=None):
        """One frame: buffers (when redrawing) + draw; returns the framebuffer."""
        if redraw:
            self.update_buffers(pointer)
        return self.draw(out=out)

    def synchronize(self) -> None:
        check(lib().isaac_hip_synchronize(self._ctx), self._ctx, "synchronize")

    def advection(self, field_idx: int = 0):
        """Host copies of field source field_idx's current advection texture (Z, Y, X)
        uint8 and of the noise texture (Z, Y, X) float32, and the time step."""
        X, Y, Z = self.local_size
        tex, noise, ts = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_int32()
        check(lib().isaac_hip_get_advection(self._ctx, int(field_idx), ctypes.byref(tex), ctypes.byref(noise),
                                            ctypes.byref(ts)), self._ctx, "get_advection")
        self.synchronize()
        a = np.zeros((Z, Y, X), np.uint8)
        n = np.zeros((Z, Y, X), np.float32)
        if tex.value:
            self._d2h(a, tex.value)
        if noise.value:
            self._d2h(n, noise.value)
        return a, n, int(ts.value)

    def video_frame_message(self, quality: int = 90) -> bytes:
        """The message serverSendFrame sends for the last rendered frame
        (lib/isaac.hpp:2845-2861): {"payload": "data:image/jpeg;base64,..."} for quality
        1..100 (libjpeg-turbo's bytes, encoded on the GPU), the raw-RGBA payload for 0."""
        cap = int(lib().isaac_hip_video_frame_capacity(self._ctx))
        buf = ctypes.create_string_buffer(cap)
        n = ctypes.c_int64(0)
        check(lib().isaac_hip_video_frame_message(self._ctx, int(quality), buf, cap, ctypes.byref(n)), self._ctx,
              "video_frame_message")
        return ctypes.string_at(buf, n.value)

    def min_max(self) -> Tuple[np.ndarray, np.ndarray]:
        """The min/max request (sendMinMax, lib/isaac.hpp:1951-2006): per source in the
        combined order, min and max of its functor-chain values (float32 arrays).
        Non-persistent sources report the values of the last update_buffers."""
        n = len(self.volume_sources) + len(self.field_sources) + len(self.particle_sources)
        mins = np.zeros(max(n, 1), np.float32)
        maxs = np.zeros(max(n, 1), np.float32)
        check(lib().isaac_hip_min_max(self._ctx, mins.ctypes.data, maxs.ctypes.data, max(n, 1)), self._ctx, "min_max")
        return mins[:n], maxs[:n]

    def timings(self) -> Dict[str, float]:
        t = _lib.Timings()
        check(lib().isaac_hip_get_timings(self._ctx, ctypes.byref(t)), self._ctx, "get_timings")
        self.kernel_time, self.copy_time = t.kernel_us, t.copy_us
        self.buffer_time, self.optimization_buffer_time = t.buffer_us, t.optimization_buffer_us
        return {name: float(getattr(t, name)) for name, _ in _lib.Timings._fields_}

    TUNE_MARCH_PREFETCH = 0
    TUNE_MARCH_WAVES = 1
    TUNE_MARCH_TILES = 2
    TUNE_TEX_LAYOUT = 3
    TUNE_MARCH_THREADS = 4
    TUNE_MARCH_OFFSETS = 5
    TUNE_TIMING_EVENTS = 6
    TUNE_FUSED_SHADING = 7
    tex_layout = 0
    render_form = 0

    def set_tuning(self, key: int, value: int) -> None:
        """TUNE_TEX_LAYOUT takes effect at the next update_buffers (a render before it
        raises)."""
        check(lib().isaac_hip_set_tuning(self._ctx, int(key), int(value)), self._ctx, "set_tuning")
        if key == self.TUNE_TEX_LAYOUT:
            self.tex_layout = int(value)

    # render forms (no counterpart in the reference; include/isaac_hip.h)
    FORM_COMBINED = 0
    FORM_BRICKS = 1
    F32, F16, BF16 = 0, 1, 2

    def set_render_form(self, form: int, brick_dtype: int = 0) -> None:
        """COMBINED: the reference's optimized renderer; BRICKS: post-classification
        march over 8^3 field bricks stored in brick_dtype (F32/F16/BF16 + FP32 scale).
        Takes effect at the next update_buffers."""
        check(lib().isaac_hip_set_render_form(self._ctx, int(form), int(brick_dtype)), self._ctx, "set_render_form")
        self.render_form = int(form)

    def bricks(self, source: int):
        """Host copies of a source's bricks: (pairs, scales) with pairs shaped
        (nbricks, 9, 9, 8, dim, 2) as float32 (F32) or uint16 bit patterns (F16/BF16)
        and scales (nbricks, dim)."""
        pp, ps = ctypes.c_void_p(), ctypes.c_void_p()
        nb, pb = ctypes.c_int64(), ctypes.c_int32()
        check(lib().isaac_hip_get_bricks(self._ctx, int(source), ctypes.byref(pp), ctypes.byref(ps), ctypes.byref(nb),
                                         ctypes.byref(pb)), self._ctx, "get_bricks")
        if not pp.value:
            raise IsaacHipError(f"source {source} has no bricks (render form BRICKS, enabled, after update_buffers)")
        src = (self.volume_sources + self.field_sources)[source]
        dim = int(src.feature_dim) if src.persistent else 1  # non-persistent: the functor-applied scalar
        self.synchronize()
        dt = np.float32 if pb.value == 8 else np.uint16
        pairs = np.zeros((nb.value, 9, 9, 8, dim, 2), dt)
        scales = np.zeros((nb.value, dim), np.float32)
        self._d2h(pairs, pp.value)
        self._d2h(scales, ps.value)
        return pairs, scales

    # learned occlusion (no counterpart in the reference; include/isaac_hip.h)
    AO_PLAIN = 0
    AO_LEARNED = 1
    ao_mode = 0

    def set_ao_mode(self, mode: int) -> None:
        check(lib().isaac_hip_set_ao_mode(self._ctx, int(mode)), self._ctx, "set_ao_mode")
        self.ao_mode = int(mode)

    def aonet_set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        check(lib().isaac_hip_aonet_set_params(self._ctx, p.ctypes.data, p.size), self._ctx, "aonet_set_params")

    def aonet_get_params(self) -> np.ndarray:
        n = int(lib().isaac_hip_aonet_param_count())
        p = np.zeros(n, np.float32)
        check(lib().isaac_hip_aonet_get_params(self._ctx, p.ctypes.data, n), self._ctx, "aonet_get_params")
        return p

    def aonet_clear_batch(self) -> None:
        check(lib().isaac_hip_aonet_clear_batch(self._ctx), self._ctx, "aonet_clear_batch")

    def aonet_add_crops(self, xy: np.ndarray, side: int) -> None:
        a = np.ascontiguousarray(xy, np.int32).reshape(-1, 2)
        check(lib().isaac_hip_aonet_add_crops(self._ctx, a.shape[0], a.ctypes.data, int(side)), self._ctx, "aonet_add_crops")

    def aonet_train_step(self, lr: float) -> float:
        loss = ctypes.c_float()
        check(lib().isaac_hip_aonet_train_step(self._ctx, float(lr), ctypes.byref(loss)), self._ctx, "aonet_train_step")
        return float(loss.value)

    def aonet_loss_grad(self, features: np.ndarray, target: np.ndarray):
        """(loss, gradient) of the network on host features (n, h, w, 8) and targets (n, h, w)."""
        f = np.ascontiguousarray(features, np.float32)
        t = np.ascontiguousarray(target, np.float32)
        n, h, w = t.shape
        g = np.zeros(int(lib().isaac_hip_aonet_param_count()), np.float32)
        loss = ctypes.c_float()
        check(lib().isaac_hip_aonet_loss_grad(self._ctx, n, h, w, f.ctypes.data, t.ctypes.data, ctypes.byref(loss), g.ctypes.data),
              self._ctx, "aonet_loss_grad")
        return float(loss.value), g

    def aonet_forward(self, features: np.ndarray) -> np.ndarray:
        """Raw network output (n, h, w) on host features (n, h, w, 8)."""
        f = np.ascontiguousarray(features, np.float32)
        n, h, w = f.shape[:3]
        out = np.zeros((n, h, w), np.float32)
        check(lib().isaac_hip_aonet_forward(self._ctx, n, h, w, f.ctypes.data, out.ctypes.data), self._ctx, "aonet_forward")
        return out

    def set_float_output(self, on: bool) -> None:
        check(lib().isaac_hip_set_float_output(self._ctx, int(on)), self._ctx, "set_float_output")

    def float_image(self) -> np.ndarray:
        """Final colour before RGBA8 quantisation, (H, W, 4) float32 (needs set_float_output)."""
        p = ctypes.c_void_p()
        check(lib().isaac_hip_get_float_image(self._ctx, ctypes.byref(p)), self._ctx, "get_float_image")
        if not p.value:
            raise IsaacHipError("float output is off (set_float_output(True) before draw)")
        W, H = self.framebuffer_size
        self.synchronize()
        a = np.zeros((H, W, 4), np.float32)
        self._d2h(a, p.value)
        return a

    def set_counting(self, on: bool) -> None:
        check(lib().isaac_hip_set_counting(self._ctx, int(on)), self._ctx, "set_counting")

    def counters(self) -> Dict[str, int]:
        out = np.zeros(len(_lib.CNT_NAMES), np.uint64)
        check(lib().isaac_hip_get_counters(self._ctx, out.ctypes.data), self._ctx, "get_counters")
        return dict(zip(_lib.CNT_NAMES, [int(v) for v in out]))

    # ---------------------------------------------------------------- introspection (tests)
    def gbuffer(self):
        """Copies of the G-buffer: colour (H,W,4) u8, depth (H,W), normal (H,W,3), ao (H,W)."""
        import torch

        W, H = self.framebuffer_size
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        check(lib().isaac_hip_get_gbuffer(self._ctx, *[ctypes.byref(p) for p in ptrs]), self._ctx, "get_gbuffer")
        self.synchronize()
        out = []
        for p, shape, dt in zip(ptrs, [(H, W, 4), (H, W), (H, W, 3), (H, W)], [np.uint8, np.float32, np.float32, np.float32]):
            a = np.zeros(shape, dt)
            self._d2h(a, p.value)
            out.append(a)
        return out

    def combined_textures(self):
        """(volume, iso) combined RGBA8 textures, each (Z, Y, X, 4) uint8."""
        X, Y, Z = self.local_size
        pt, nb = ctypes.c_void_p(), ctypes.c_int32()
        check(lib().isaac_hip_get_combined_textures(self._ctx, ctypes.byref(pt), ctypes.byref(nb)), self._ctx, "textures")
        self.synchronize()
        tex = np.zeros((Z, Y, X, int(nb.value)), np.uint8)
        self._d2h(tex, pt.value)
        return np.ascontiguousarray(tex[..., 0:4]), np.ascontiguousarray(tex[..., 4:8])

    def cell_list(self, idx: int):
        gx, gy, gz = self.local_particle_size
        nc = gx * gy * gz
        n = int(self.particle_sources[idx].pos.shape[0])
        pc, po, pp = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib().isaac_hip_get_cell_list(self._ctx, idx, ctypes.byref(pc), ctypes.byref(po), ctypes.byref(pp)), self._ctx, "cell_list")
        self.synchronize()
        counts = np.zeros(nc, np.int32)
        offs = np.zeros(nc + 1, np.int32)
        perm = np.zeros(n, np.int32)
        self._d2h(counts, pc.value)
        self._d2h(offs, po.value)
        self._d2h(perm, pp.value)
        return counts, offs, perm

    def _d2h(self, arr: np.ndarray, dev_ptr: int) -> None:
        check(lib().isaac_hip_copy_to_host(self._ctx, arr.ctypes.data, ctypes.c_void_p(dev_ptr), arr.nbytes), self._ctx, "copy_to_host")


# --------------------------------------------------------------------------- scene -> visualization
def from_scene(spec, device: int = 0) -> IsaacVisualization:
    """Build a visualization for a scenes.SceneSpec with its sources uploaded to `device`."""
    import torch

    dev = torch.device("cuda", device)
    vols = [VolumeSource(v.name, torch.from_numpy(np.ascontiguousarray(v.data)).to(dev), v.persistent) for v in spec.volumes]
    fields = [VolumeSource(v.name, torch.from_numpy(np.ascontiguousarray(v.data)).to(dev), v.persistent)
              for v in getattr(spec, "fields", [])]
    parts = []
    for p in spec.particles:
        rad = torch.from_numpy(p.radius).to(dev) if p.radius is not None else None
        parts.append(
            ParticleSource(p.name, torch.from_numpy(p.pos).to(dev), torch.from_numpy(p.attr).to(dev), rad, p.constant_radius)
        )
    torch.cuda.synchronize(dev)
    vis = IsaacVisualization(
        spec.name, spec.framebuffer, spec.global_size, spec.local_size, spec.particle_grid, spec.position,
        vols, fields, parts, spec.scale, spec.transfer_size, device,
    )
    apply_scene(vis, spec)
    return vis


def apply_scene(vis: IsaacVisualization, spec) -> None:
    fields = list(getattr(spec, "fields", []))
    srcs = list(spec.volumes) + fields + list(spec.particles)
    nvf = len(spec.volumes) + len(fields)
    vis.functions = [s.chain for s in srcs]
    vis.update_functions()
    for i, s in enumerate(srcs):
        vis.set_transfer_points(i, s.tf_points)
        vis.source_weight[i] = s.weight
        vis.iso_threshold[i] = getattr(s, "iso", 0.0) if i < nvf else 0.0
    adv = getattr(spec, "advection", None)
    if adv is not None:
        (vis.advection_step, vis.advection_weight, vis.advection_seeding_period,
         vis.advection_seeding_duration, vis.seed_points) = adv
    vis.update_transfer()
    vis.step = spec.step
    vis.interpolation = spec.interpolation
    vis.dither_mode = spec.dither_mode
    vis.render_mode = spec.render_mode
    vis.ambient_occlusion = {"isEnabled": spec.ao_enabled, "weight": spec.ao_weight}
    vis.background_color = list(spec.background)
    vis.viewport = spec.viewport
    while len(vis.clipping):
        vis.remove_clipping(0)
    for pos, nrm in spec.clipping:
        vis.add_clipping(*pos, *nrm)
    cam = spec.camera
    vis.look_at = np.asarray(cam.look_at, np.float64)
    vis.rotation = np.asarray(cam.rotation, np.float64)
    vis.distance = float(cam.distance)
    vis._update_modelview()
