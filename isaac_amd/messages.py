"""The master rank's JSON messages to the ISAAC server through the C-ABI
(isaac_hip_register_message, isaac_hip_period_message, isaac_hip_exit_message,
isaac_hip_server_send_bytes): the register message of the constructor
(lib/isaac.hpp:834-892), the periodic state message of each step
(lib/isaac.hpp:2546-2838), the exit message (lib/isaac.hpp:2110-2118) and
serverSend's bytes (lib/isaac/isaac_communicator.hpp:125-190).  Host only."""
from __future__ import annotations

import ctypes
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from ._lib import IsaacHipError, SteeringState, lib

SEND = [
    "projection", "look_at", "rotation", "distance", "transfer", "functions", "weight", "interpolation",
    "render_optimization", "step", "seed_points", "advection_step", "advection_weight", "advection_seeding_period",
    "advection_seeding_duration", "advection_on_pause", "advection_border", "iso_threshold", "min_max", "background",
    "clipping", "ao", "render_mode", "dither_mode", "init",
]
FLAG = {name: 1 << i for i, name in enumerate(SEND)}


def send_mask(names: Iterable[str]) -> int:
    m = 0
    for n in names:
        if n not in FLAG:
            raise ValueError(f"unknown section {n!r}")
        m |= FLAG[n]
    return m


def _call(fn, *args) -> str:
    n = ctypes.c_int64(0)
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        rc = fn(*args, buf, cap, ctypes.byref(n))
        if rc == 0:
            return buf.raw[: n.value].decode("utf-8")
        if rc == -1 and n.value + 1 > cap:
            cap = n.value + 1
            continue
        raise IsaacHipError(f"message: error {rc}")


class _Arrays:
    """Keeps the ctypes arrays a SteeringState points to alive."""

    def __init__(self):
        self.keep = []

    def strings(self, items: Sequence[str]):
        arr = (ctypes.c_char_p * max(1, len(items)))(*[s.encode("utf-8") for s in items])
        self.keep.append(arr)
        return ctypes.cast(arr, ctypes.POINTER(ctypes.c_char_p))

    def array(self, ctype, values, n_min=1):
        a = np.ascontiguousarray(values if len(values) else np.zeros(n_min), dtype=ctype)
        self.keep.append(a)
        return a


def steering_state(s) -> Tuple[SteeringState, _Arrays]:
    """A SteeringState from an object with the fields of oracle-style State
    (source_names, feature_dims, n_volume_field, functions, ...)."""
    keep = _Arrays()
    st = SteeringState()
    n = len(s.source_names)
    st.n_sources = n
    st.n_volume_field = int(s.n_volume_field)
    st.source_names = keep.strings(list(s.source_names))
    fd = keep.array(np.int32, list(s.feature_dims))
    st.feature_dims = fd.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    st.functions = keep.strings(list(s.functions) or [""] * n)
    fe = keep.array(np.int32, list(s.function_errors) or [0] * n)
    st.function_errors = fe.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    st.transfer_size = int(s.transfer_size)
    if s.transfer_lut is not None:
        lut = keep.array(np.float32, np.asarray(s.transfer_lut, np.float32).reshape(-1))
        st.transfer_lut = lut.ctypes.data
    counts = keep.array(np.int32, [len(p) for p in s.transfer_points] or [0] * n)
    st.transfer_point_counts = counts.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    vals = keep.array(np.uint32, [v for p in s.transfer_points for v, _ in p])
    st.transfer_point_values = vals.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
    rgba = keep.array(np.float32, [c for p in s.transfer_points for _, col in p for c in col])
    st.transfer_point_rgba = rgba.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    w = keep.array(np.float32, list(s.weights))
    st.weights = w.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    iso = keep.array(np.float32, list(s.iso_thresholds))
    st.iso_thresholds = iso.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    mm = keep.array(np.float32, [x for pair in s.min_max for x in pair])
    st.min_max = mm.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    st.step = float(s.step)
    st.interpolation = int(bool(s.interpolation))
    st.render_optimization = int(bool(s.render_optimization))
    st.seed_points = int(s.seed_points)
    st.advection_step = float(s.advection_step)
    st.advection_weight = float(s.advection_weight)
    st.advection_seeding_period = int(s.advection_seeding_period)
    st.advection_seeding_duration = int(s.advection_seeding_duration)
    st.advection_on_pause = int(bool(s.advection_on_pause))
    st.advection_border = int(bool(s.advection_border))
    st.background[:] = [float(x) for x in s.background]
    st.n_clipping = len(s.clipping)
    cp = keep.array(np.float32, [x for p, _ in s.clipping for x in p])
    cn = keep.array(np.float32, [x for _, nrm in s.clipping for x in nrm])
    st.clipping_position = cp.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    st.clipping_normal = cn.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    st.ao_enabled = int(bool(s.ao_enabled))
    st.ao_weight = float(s.ao_weight)
    st.render_mode = int(s.render_mode)
    st.dither_mode = int(s.dither_mode)
    st.projection[:] = [float(x) for x in s.projection]
    st.look_at[:] = [float(x) for x in s.look_at]
    st.rotation[:] = [float(x) for x in s.rotation]
    st.distance = float(s.distance)
    meta = (s.metadata or "{}").encode("utf-8")
    keep.keep.append(meta)
    st.metadata = meta
    return st, keep


class Session:
    """The master rank's message state (jsonInitRoot) across steps."""

    def __init__(self):
        self._s = ctypes.c_void_p()
        if lib().isaac_hip_session_create(ctypes.byref(self._s)) != 0:
            raise IsaacHipError("session_create failed")

    def close(self) -> None:
        if self._s:
            lib().isaac_hip_session_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def register(self, state, name: str, nodes: int, framebuffer, global_size_scaled) -> str:
        st, keep = steering_state(state)
        fb = (ctypes.c_int32 * 2)(*[int(x) for x in framebuffer])
        gs = (ctypes.c_int32 * 3)(*[int(x) for x in global_size_scaled])
        return _call(lib().isaac_hip_register_message, self._s, ctypes.byref(st), name.encode("utf-8"), int(nodes), fb, gs)

    def period(self, state, meta_nr: int, send: Iterable[str]) -> str:
        st, keep = steering_state(state)
        return _call(lib().isaac_hip_period_message, self._s, ctypes.byref(st), int(meta_nr), send_mask(send))


def exit_message() -> str:
    return _call(lib().isaac_hip_exit_message)


def server_send_bytes(uid: int, content: Optional[str], finishing: bool) -> str:
    c = content.encode("utf-8") if content is not None else None
    return _call(lib().isaac_hip_server_send_bytes, int(uid), c, int(bool(finishing)))


__all__ = ["FLAG", "SEND", "Session", "exit_message", "send_mask", "server_send_bytes", "steering_state"]
