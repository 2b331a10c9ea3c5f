"""CPU ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes front end of oracle/liboracle.so (the C++ restatement in
oracle/isaac_oracle.hpp) plus independent Python restatements of the
host-side steps of the reference's render path:

  build_lut      updateTransfer            lib/isaac.hpp:1244-1266
  parse_chain    updateFunctions parser    lib/isaac.hpp:1079-1187
  perspective    DefaultController          lib/isaac/isaac_controllers.hpp:33-37 (glm perspectiveRH_NO)
  modelview      updateModelview            lib/isaac.hpp:2878-2889
  clip_normal    editClipping               lib/isaac.hpp:958-986

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module; the product never does.
"""
from __future__ import annotations

import ctypes
import math
import os
import re
import subprocess
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_SOURCES = 16
MAX_FUNCTORS = 3
MAX_CLIPPING = 16
FUNCTORS = ["idem", "add", "mul", "length", "pow", "sum"]  # lib/isaac/isaac_functors.hpp:185-208
Z_NEAR = float(np.float32(0.002))  # ISAAC_Z_NEAR is a float literal (lib/isaac/isaac_defines.hpp:53)
Z_FAR = 100.0


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-C", HERE])
    return LIB_PATH


_lib = None


def _bind(path: str):
    L = ctypes.CDLL(path)
    L.oracle_render_f32.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_render_f64.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.oracle_apply_chain.restype = ctypes.c_float
    L.oracle_ray.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    L.oracle_perspective.argtypes = [ctypes.c_double] * 4 + [ctypes.c_void_p]
    return L


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = _bind(LIB_PATH)
    return _lib


def build_mutant(mutant: int, out_dir: str) -> str:
    """Compile the restatement with ORACLE_MUTANT=<mutant> (one planted mistake, see
    isaac_oracle.hpp) into out_dir; used only by the tests that show the closed-form
    checks reject each planted mistake."""
    path = os.path.join(out_dir, f"liboracle_mutant{int(mutant)}.so")
    subprocess.check_call(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                           "-shared", f"-DORACLE_MUTANT={int(mutant)}", "-o", path, os.path.join(HERE, "oracle_capi.cpp")])
    return path


class use_library:
    """Context manager: route this module's calls through another build of the
    restatement (a mutant from build_mutant) and restore the shipped one after."""

    def __init__(self, path: str):
        self.path = path

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = _bind(self.path)
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


# --------------------------------------------------------------------------- host-side restatements
def build_lut(points: Sequence[Tuple[int, Sequence[float]]], T: int) -> np.ndarray:
    """updateTransfer: segments [p_k, p_{k+1}) of (c_k*(w-j) + c_{k+1}*j)/w in
    float32; std::map insert keeps the first value of a duplicate key; entries
    never written stay zero."""
    d = {}
    for k, rgba in points:
        k = int(k) & 0xFFFFFFFF
        if k not in d:
            d[k] = np.asarray(rgba, dtype=np.float32)
    keys = sorted(d)
    lut = np.zeros((T, 4), dtype=np.float32)
    for a, b in zip(keys, keys[1:]):
        width = b - a
        n = max(0, min(width, T - a))
        if n == 0:
            continue
        j = np.arange(n, dtype=np.int64)
        wj = (width - j).astype(np.float32)[:, None]
        jj = j.astype(np.float32)[:, None]
        lut[a : a + n] = (d[a][None, :] * wj + d[b][None, :] * jj) / np.float32(width)
    return lut


_FLOAT_PREFIX = re.compile(
    r"^[ \t\n\v\f\r]*[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan)", re.IGNORECASE
)


def _stof(s: str) -> float:
    m = _FLOAT_PREFIX.match(s)
    if not m:
        raise ValueError(s)
    v = float(m.group(0))
    if math.isfinite(v) and abs(v) > float(np.finfo(np.float32).max):
        raise OverflowError(s)
    return float(np.float32(v))


def parse_chain(src: str) -> Tuple[List[int], List[List[float]], int]:
    """updateFunctions' parser for one source: returns (bytecode[3],
    params[3][4], errorCode)."""
    bytecode = [0, 0, 0]
    params = [[0.0] * 4 for _ in range(MAX_FUNCTORS)]
    error = 0
    source = src
    elem = 0
    again = True
    while again:
        pos = source.find("|")
        if pos < 0:
            again = False
            token = source
        else:
            token = source[:pos]
            source = source[pos + 1 :]
        if elem >= MAX_FUNCTORS:
            error = 1
            break
        token = "".join(ch for ch in token if ch not in " \t\n\v\f\r")
        t_begin = token.find("(")
        if t_begin < 0:
            params[elem] = [0.0] * 4
        else:
            t_end = token.find(")")
            if t_end < 0:
                error = -1
                break
            if t_end - t_begin == 1:
                params[elem] = [0.0] * 4
            else:
                parameters = token[t_begin + 1 : t_end]
                arr = params[elem]
                p_elem = 0
                p_again = True
                while p_again:
                    ppos = parameters.find(",")
                    if ppos < 0:
                        p_again = False
                        par = parameters
                    else:
                        par = parameters[:ppos]
                        parameters = parameters[ppos + 1 :]
                    if p_elem >= 4:
                        error = 2
                        break
                    try:
                        arr[p_elem] = _stof(par)
                    except (ValueError, OverflowError):
                        error = -2
                        p_elem += 1
                        break
                    p_elem += 1
                for k in range(p_elem, 4):
                    arr[k] = arr[k - 1]
        if t_begin >= 0:
            token = token[:t_begin]
        if token in FUNCTORS:
            bytecode[elem] = FUNCTORS.index(token)
        else:
            error = -2
            break
        elem += 1
    for k in range(elem, MAX_FUNCTORS):
        bytecode[k] = 0
        params[k] = [0.0] * 4
    return bytecode, params, error


def perspective(fovy: float, aspect: float, zn: float, zf: float) -> np.ndarray:
    """glm perspectiveRH_NO, column-major 16 doubles."""
    t = math.tan(fovy / 2.0)
    m = np.zeros(16)
    m[0] = 1.0 / (aspect * t)
    m[5] = 1.0 / t
    m[10] = -(zf + zn) / (zf - zn)
    m[11] = -1.0
    m[14] = -(2.0 * zf * zn) / (zf - zn)
    return m


def default_projection(w: int, h: int) -> np.ndarray:
    # DefaultController passes 45.0 to glm::perspective (radians in glm 0.9.9)
    return perspective(45.0, float(w) / float(h), Z_NEAR, Z_FAR)


def modelview(look_at, rotation: np.ndarray, distance: float) -> np.ndarray:
    """distanceM * rotationM * translationM (lib/isaac.hpp:2878-2889), returned
    column-major (glm value_ptr order).  `rotation` is the math matrix R."""
    T = np.eye(4)
    T[:3, 3] = np.asarray(look_at, dtype=np.float64)
    R = np.eye(4)
    R[:3, :3] = np.asarray(rotation, dtype=np.float64)
    D = np.eye(4)
    D[2, 3] = distance
    M = D @ R @ T
    return M.T.reshape(16).copy()


def clip_normal(n: Sequence[float], scale: Sequence[float]) -> Optional[np.ndarray]:
    f = np.float32
    ns = np.array([f(n[0]) * f(scale[0]), f(n[1]) * f(scale[1]), f(n[2]) * f(scale[2])], dtype=np.float32)
    l = np.sqrt(f(f(ns[0] * ns[0]) + f(ns[1] * ns[1])) + f(ns[2] * ns[2]))
    if l == 0:
        return None
    return (ns / l).astype(np.float32)


# --------------------------------------------------------------------------- ctypes structs
class OracleScene(ctypes.Structure):
    _fields_ = [
        ("global_size", ctypes.c_int64 * 3),
        ("local_size", ctypes.c_int64 * 3),
        ("particle_grid", ctypes.c_int64 * 3),
        ("position", ctypes.c_int32 * 3),
        ("scale", ctypes.c_float * 3),
        ("fb_w", ctypes.c_int32),
        ("fb_h", ctypes.c_int32),
        ("viewport", ctypes.c_int32 * 4),
        ("transfer_size", ctypes.c_int32),
        ("n_volume", ctypes.c_int32),
        ("n_particle", ctypes.c_int32),
        ("vol_data", ctypes.c_void_p * MAX_SOURCES),
        ("vol_dim", ctypes.c_int32 * MAX_SOURCES),
        ("vol_stride", (ctypes.c_int64 * 3) * MAX_SOURCES),
        ("part_pos", ctypes.c_void_p * MAX_SOURCES),
        ("part_attr", ctypes.c_void_p * MAX_SOURCES),
        ("part_dim", ctypes.c_int32 * MAX_SOURCES),
        ("part_radius", ctypes.c_void_p * MAX_SOURCES),
        ("part_const_radius", ctypes.c_float * MAX_SOURCES),
        ("part_count", ctypes.c_int64 * MAX_SOURCES),
        ("lut", ctypes.c_void_p * (2 * MAX_SOURCES)),
        ("bytecode", (ctypes.c_int32 * MAX_FUNCTORS) * (2 * MAX_SOURCES)),
        ("params", ((ctypes.c_float * 4) * MAX_FUNCTORS) * (2 * MAX_SOURCES)),
        ("weight", ctypes.c_float * (2 * MAX_SOURCES)),
        ("iso", ctypes.c_float * (2 * MAX_SOURCES)),
        ("step", ctypes.c_float),
        ("interpolation", ctypes.c_int32),
        ("dither_mode", ctypes.c_int32),
        ("ao_enabled", ctypes.c_int32),
        ("ao_weight", ctypes.c_float),
        ("render_mode", ctypes.c_int32),
        ("n_clip", ctypes.c_int32),
        ("clip_pos", (ctypes.c_float * 3) * MAX_CLIPPING),
        ("clip_normal", (ctypes.c_float * 3) * MAX_CLIPPING),
        ("proj", ctypes.c_double * 16),
        ("mv", ctypes.c_double * 16),
        ("bg", ctypes.c_float * 4),
        ("threads", ctypes.c_int32),
        ("vol_nonpersistent", ctypes.c_int32 * MAX_SOURCES),
        ("vol_adv", ctypes.c_void_p * MAX_SOURCES),
    ]


class OraclePostOutputs(ctypes.Structure):
    _fields_ = [
        ("rgba", ctypes.c_void_p),
        ("rgba_f", ctypes.c_void_p),
        ("pre", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("ao", ctypes.c_void_p),
        ("counters", ctypes.c_uint64 * 8),
        ("seconds", ctypes.c_double),
    ]


class OracleOutputs(ctypes.Structure):
    _fields_ = [
        ("rgba", ctypes.c_void_p),
        ("solid", ctypes.c_void_p),
        ("pre", ctypes.c_void_p),
        ("pre_depth", ctypes.c_void_p),
        ("depth", ctypes.c_void_p),
        ("normal", ctypes.c_void_p),
        ("ao", ctypes.c_void_p),
        ("vol_tex", ctypes.c_void_p),
        ("iso_tex", ctypes.c_void_p),
        ("counters", ctypes.c_uint64 * 8),
        ("seconds_classify", ctypes.c_double),
        ("seconds_render", ctypes.c_double),
        ("seconds_bin", ctypes.c_double),
    ]


COUNTER_NAMES = ["vol_samples", "iso_samples", "iso_hits", "cells_visited", "sphere_tests", "vol_rays", "iso_rays", "part_rays"]


class PreparedScene:
    """A SceneSpec converted to the oracle's flat struct (keeps numpy buffers alive)."""

    def __init__(self, spec, proj: Optional[np.ndarray] = None, mv: Optional[np.ndarray] = None, threads: int = 0):
        self.spec = spec
        self.keep = []
        sc = OracleScene()
        sc.global_size[:] = spec.global_size
        sc.local_size[:] = spec.local_size
        sc.particle_grid[:] = spec.particle_grid
        sc.position[:] = spec.position
        sc.scale[:] = spec.scale
        W, H = spec.framebuffer
        sc.fb_w, sc.fb_h = W, H
        sc.viewport[:] = spec.vp()
        T = spec.transfer_size
        sc.transfer_size = T
        vols = list(spec.volumes) + list(getattr(spec, "fields", []))
        sc.n_volume = len(vols)
        sc.n_particle = len(spec.particles)
        nv = len(vols)
        sources = vols + list(spec.particles)
        for s, v in enumerate(vols):
            data = np.ascontiguousarray(v.data, dtype=np.float32)
            self.keep.append(data)
            sc.vol_data[s] = data.ctypes.data
            sc.vol_dim[s] = v.feature_dim
            X, Y, Z = spec.local_size
            sc.vol_stride[s][0], sc.vol_stride[s][1], sc.vol_stride[s][2] = 1, X, X * Y
            sc.vol_nonpersistent[s] = 0 if v.persistent else 1
        for s, p in enumerate(spec.particles):
            pos = np.ascontiguousarray(p.pos, dtype=np.float32)
            attr = np.ascontiguousarray(p.attr, dtype=np.float32)
            self.keep += [pos, attr]
            sc.part_pos[s] = pos.ctypes.data
            sc.part_attr[s] = attr.ctypes.data
            sc.part_dim[s] = p.feature_dim
            if p.radius is not None:
                r = np.ascontiguousarray(p.radius, dtype=np.float32)
                self.keep.append(r)
                sc.part_radius[s] = r.ctypes.data
            else:
                sc.part_radius[s] = None
            sc.part_const_radius[s] = p.constant_radius
            sc.part_count[s] = pos.shape[0]
        self.luts = []
        for s, src in enumerate(sources):
            lut = build_lut(src.tf_points, T)
            self.luts.append(lut)
            sc.lut[s] = lut.ctypes.data
            bc, pr, err = parse_chain(src.chain)
            for k in range(MAX_FUNCTORS):
                sc.bytecode[s][k] = bc[k]
                for c in range(4):
                    sc.params[s][k][c] = pr[k][c]
            sc.weight[s] = src.weight
            sc.iso[s] = getattr(src, "iso", 0.0) if s < nv else 0.0
        sc.step = max(spec.step, float(np.float32(0.01)))
        sc.interpolation = int(spec.interpolation)
        sc.dither_mode = spec.dither_mode
        sc.ao_enabled = int(spec.ao_enabled)
        sc.ao_weight = spec.ao_weight
        sc.render_mode = spec.render_mode
        n = 0
        for pos, nrm in spec.clipping:
            cn = clip_normal(nrm, spec.scale)
            if cn is None or n >= MAX_CLIPPING:
                continue
            sc.clip_pos[n][:] = pos
            sc.clip_normal[n][:] = cn
            n += 1
        sc.n_clip = n
        cam = spec.camera
        self.proj = default_projection(W, H) if proj is None else np.asarray(proj, dtype=np.float64)
        self.mv = modelview(cam.look_at, cam.rotation, cam.distance) if mv is None else np.asarray(mv, dtype=np.float64)
        sc.proj[:] = self.proj.tolist()
        sc.mv[:] = self.mv.tolist()
        sc.bg[:] = spec.background
        sc.threads = threads
        self.sc = sc


class OracleResult:
    pass


def render(spec, precision: int = 32, threads: int = 0, want_textures: bool = False, prepared: Optional[PreparedScene] = None):
    """Render one frame (classification + binning + drawCallBack) on the CPU."""
    ps = prepared or PreparedScene(spec, threads=threads)
    ps.sc.threads = threads
    W, H = spec.framebuffer
    X, Y, Z = spec.local_size
    res = OracleResult()
    res.rgba = np.zeros((H, W, 4), np.uint8)
    res.solid = np.zeros((H, W, 4), np.uint8)
    res.pre = np.zeros((H, W, 4), np.uint8)
    res.pre_depth = np.zeros((H, W), np.float32)
    res.depth = np.zeros((H, W), np.float32)
    res.normal = np.zeros((H, W, 3), np.float32)
    res.ao = np.zeros((H, W), np.float32)
    out = OracleOutputs()
    out.rgba = res.rgba.ctypes.data
    out.solid = res.solid.ctypes.data
    out.pre = res.pre.ctypes.data
    out.pre_depth = res.pre_depth.ctypes.data
    out.depth = res.depth.ctypes.data
    out.normal = res.normal.ctypes.data
    out.ao = res.ao.ctypes.data
    if want_textures:
        res.vol_tex = np.zeros((Z, Y, X, 4), np.uint8)
        res.iso_tex = np.zeros((Z, Y, X, 4), np.uint8)
        out.vol_tex = res.vol_tex.ctypes.data
        out.iso_tex = res.iso_tex.ctypes.data
    fn = lib().oracle_render_f32 if precision == 32 else lib().oracle_render_f64
    rc = fn(ctypes.byref(ps.sc), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"oracle render failed rc={rc}")
    res.counters = dict(zip(COUNTER_NAMES, [int(c) for c in out.counters]))
    res.seconds = {"classify": out.seconds_classify, "bin": out.seconds_bin, "render": out.seconds_render}
    res.prepared = ps
    return res


def set_advection(ps: PreparedScene, textures: dict) -> None:
    """Attach advection textures {source index: (Z, Y, X) uint8} of field sources to a
    prepared scene (the classification multiplies their alpha by texture / 255)."""
    for s, t in textures.items():
        a = np.ascontiguousarray(t, np.uint8)
        ps.keep.append(a)
        ps.sc.vol_adv[s] = a.ctypes.data


def noise(local_size, scale, seed_points: int) -> np.ndarray:
    """updateNoiseTexture: (Z, Y, X) float32."""
    X, Y, Z = local_size
    out = np.zeros((Z, Y, X), np.float32)
    loc = (ctypes.c_int64 * 3)(*local_size)
    sc = (ctypes.c_float * 3)(*scale)
    lib().oracle_noise(loc, sc, ctypes.c_uint32(seed_points), out.ctypes.data_as(ctypes.c_void_p))
    return out


def advect(field: np.ndarray, noise_tex: np.ndarray, history: np.ndarray, scale, step_size: float,
           history_weight: float, seeding: bool) -> np.ndarray:
    """One GenerateAdvectionTextureKernel pass: field (Z, Y, X, 3) float32 -> (Z, Y, X) uint8."""
    Z, Y, X = history.shape
    f = np.ascontiguousarray(field, np.float32)
    nz = np.ascontiguousarray(noise_tex, np.float32)
    h = np.ascontiguousarray(history, np.uint8)
    out = np.zeros((Z, Y, X), np.uint8)
    loc = (ctypes.c_int64 * 3)(X, Y, Z)
    sc = (ctypes.c_float * 3)(*scale)
    lib().oracle_advect(f.ctypes.data_as(ctypes.c_void_p), loc, sc, ctypes.c_float(step_size), ctypes.c_float(history_weight),
                        ctypes.c_int(int(seeding)), nz.ctypes.data_as(ctypes.c_void_p), h.ctypes.data_as(ctypes.c_void_p),
                        out.ctypes.data_as(ctypes.c_void_p))
    return out


def min_max(spec, prepared: Optional[PreparedScene] = None) -> Tuple[np.ndarray, np.ndarray]:
    """The min/max request: per source (volume sources first, then particle sources)
    the min and max of its functor-chain values, float32 (oracle_min_max)."""
    ps = prepared or PreparedScene(spec)
    n = int(ps.sc.n_volume) + int(ps.sc.n_particle)
    mins = np.zeros(max(n, 1), np.float32)
    maxs = np.zeros(max(n, 1), np.float32)
    rc = lib().oracle_min_max(ctypes.byref(ps.sc), mins.ctypes.data_as(ctypes.c_void_p), maxs.ctypes.data_as(ctypes.c_void_p))
    if rc != 0:
        raise RuntimeError(f"oracle min_max failed rc={rc}")
    return mins[:n], maxs[:n]


BRICK_RAW, BRICK_F32, BRICK_F16, BRICK_BF16 = -1, 0, 1, 2


def render_post(spec, precision: int = 64, dtype: int = BRICK_RAW, stride: int = 1, threads: int = 0,
                prepared: Optional[PreparedScene] = None):
    """One frame of the brick (post-classification) form on the CPU.

    precision 32: ray setup in binary64, sampling/compositing in binary32 (the
    library's arithmetic); 64: everything in binary64.  dtype BRICK_RAW samples the
    float sources (the FP64 reference image), BRICK_F32/F16/BF16 the bricks.
    stride > 1 renders only the pixels of sampled_mask(W, H, stride, SSAO on): with
    SSAO on the stride must divide 10 and the viewport be the whole framebuffer."""
    ps = prepared or PreparedScene(spec, threads=threads)
    ps.sc.threads = threads
    W, H = spec.framebuffer
    res = OracleResult()
    res.rgba = np.zeros((H, W, 4), np.uint8)
    res.rgba_f = np.zeros((H, W, 4), np.float32)
    res.pre = np.zeros((H, W, 4), np.uint8)
    res.depth = np.zeros((H, W), np.float32)
    res.normal = np.zeros((H, W, 3), np.float32)
    res.ao = np.zeros((H, W), np.float32)
    out = OraclePostOutputs()
    out.rgba = res.rgba.ctypes.data
    out.rgba_f = res.rgba_f.ctypes.data
    out.pre = res.pre.ctypes.data
    out.depth = res.depth.ctypes.data
    out.normal = res.normal.ctypes.data
    out.ao = res.ao.ctypes.data
    rc = lib().oracle_render_post(ctypes.byref(ps.sc), int(precision), int(dtype), int(stride), ctypes.byref(out))
    if rc != 0:
        raise RuntimeError(f"oracle post render failed rc={rc}")
    res.counters = dict(zip(COUNTER_NAMES, [int(c) for c in out.counters]))
    res.seconds = out.seconds
    res.prepared = ps
    return res


def sampled_mask(width: int, height: int, stride: int, ssao: bool) -> np.ndarray:
    """(H, W) bool: the pixels render_post renders at `stride`.  x % stride == 0 and
    y % stride == 0; with SSAO on also (W-1-x) % stride == 0 and (H-1-y) % stride == 0,
    which closes the set under the SSAO taps x + 10i clamped to [0, W-1]."""
    def kept(n):
        v = np.arange(n)
        k = v % stride == 0
        if ssao:
            k |= (n - 1 - v) % stride == 0
        return k

    return kept(height)[:, None] & kept(width)[None, :]


def brick_scales(data: np.ndarray, dtype: int) -> np.ndarray:
    """Per-brick, per-component scales of a (Z, Y, X[, C]) float32 source."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    Z, Y, X = data.shape[:3]
    dim = 1 if data.ndim == 3 else data.shape[3]
    nb = ((X + 7) // 8) * ((Y + 7) // 8) * ((Z + 7) // 8)
    out = np.zeros((nb, dim), np.float32)
    st = (ctypes.c_int64 * 3)(1, X, X * Y)
    loc = (ctypes.c_int64 * 3)(X, Y, Z)
    lib().oracle_brick_scales(ctypes.c_void_p(data.ctypes.data), st, dim, loc, int(dtype), ctypes.c_void_p(out.ctypes.data))
    return out


def sample_field(data: np.ndarray, dtype: int, pos: np.ndarray, linear: bool = True, precision: int = 64) -> np.ndarray:
    """Brick-form samples (n, 4) of a (Z, Y, X[, C]) float32 source at positions (n, 3)."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    Z, Y, X = data.shape[:3]
    dim = 1 if data.ndim == 3 else data.shape[3]
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros((pos.shape[0], 4), np.float64)
    st = (ctypes.c_int64 * 3)(1, X, X * Y)
    loc = (ctypes.c_int64 * 3)(X, Y, Z)
    lib().oracle_sample_field(ctypes.c_void_p(data.ctypes.data), st, dim, loc, int(dtype), int(precision), int(linear),
                              ctypes.c_int64(pos.shape[0]), ctypes.c_void_p(pos.ctypes.data), ctypes.c_void_p(out.ctypes.data))
    return out


def f32_to_f16_bits(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib().oracle_f32_to_f16(ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(out.ctypes.data), ctypes.c_int64(x.size))
    return out


def f16_bits_to_f32(h: np.ndarray) -> np.ndarray:
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.zeros(h.shape, np.float32)
    lib().oracle_f16_to_f32(ctypes.c_void_p(h.ctypes.data), ctypes.c_void_p(out.ctypes.data), ctypes.c_int64(h.size))
    return out


def f32_to_bf16_bits(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.zeros(x.shape, np.uint16)
    lib().oracle_f32_to_bf16(ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(out.ctypes.data), ctypes.c_int64(x.size))
    return out


def psnr_float(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR of two [0, 1] float images (peak 1)."""
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0.0 else 10.0 * math.log10(1.0 / mse)


def bin_particles(pos: np.ndarray, local_size, grid):
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    n = pos.shape[0]
    ncell = int(grid[0] * grid[1] * grid[2])
    cell_of = np.zeros(n, np.int32)
    counts = np.zeros(ncell, np.int32)
    offsets = np.zeros(ncell + 1, np.int64)
    perm = np.zeros(n, np.int64)
    rel = np.zeros((n, 3), np.float32)
    L = (ctypes.c_int64 * 3)(*local_size)
    G = (ctypes.c_int64 * 3)(*grid)
    lib().oracle_bin_particles(
        ctypes.c_void_p(pos.ctypes.data), ctypes.c_int64(n), L, G,
        ctypes.c_void_p(cell_of.ctypes.data), ctypes.c_void_p(counts.ctypes.data),
        ctypes.c_void_p(offsets.ctypes.data), ctypes.c_void_p(perm.ctypes.data), ctypes.c_void_p(rel.ctypes.data),
    )
    return cell_of, counts, offsets, perm, rel


def psnr(a: np.ndarray, b: np.ndarray, peak: float = 255.0) -> float:
    d = a.astype(np.float64) - b.astype(np.float64)
    mse = float(np.mean(d * d))
    if mse == 0.0:
        return float("inf")
    return 10.0 * math.log10(peak * peak / mse)
