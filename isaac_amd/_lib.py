"""ctypes binding of include/isaac_hip.h (isaac_amd/libisaac_hip.so).

Loading fails loudly when the library has not been built: there is no CPU or
PyTorch fallback behind this module.
"""
from __future__ import annotations

import ctypes
import os
import re
from typing import List

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libisaac_hip.so")
HEADER = os.path.join(ROOT, "include", "isaac_hip.h")

MAX_SOURCES = 16
CNT_NAMES = ["vol_samples", "iso_samples", "iso_hits", "cells_visited", "sphere_tests", "vol_rays", "iso_rays", "part_rays",
             "march_steps", "march_wave_steps"]

ERRORS = {0: "ok", -1: "bad argument", -2: "device error", -3: "bad state", -4: "unsupported", -5: "allocation failed"}


class IsaacHipError(RuntimeError):
    pass


class Config(ctypes.Structure):
    _fields_ = [
        ("framebuffer_size", ctypes.c_int32 * 2),
        ("global_size", ctypes.c_int64 * 3),
        ("local_size", ctypes.c_int64 * 3),
        ("local_particle_size", ctypes.c_int64 * 3),
        ("position", ctypes.c_int32 * 3),
        ("scale", ctypes.c_float * 3),
        ("transfer_size", ctypes.c_int32),
        ("n_volume_sources", ctypes.c_int32),
        ("n_field_sources", ctypes.c_int32),
        ("n_particle_sources", ctypes.c_int32),
    ]


class VolumeDesc(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("dtype", ctypes.c_int32),
        ("feature_dim", ctypes.c_int32),
        ("stride", ctypes.c_int64 * 3),
        ("guard_size", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
    ]


class ParticleDesc(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_void_p),
        ("attribute", ctypes.c_void_p),
        ("feature_dim", ctypes.c_int32),
        ("radius", ctypes.c_void_p),
        ("constant_radius", ctypes.c_float),
        ("count", ctypes.c_int64),
    ]


class Timings(ctypes.Structure):
    _fields_ = [
        ("kernel_us", ctypes.c_double),
        ("copy_us", ctypes.c_double),
        ("buffer_us", ctypes.c_double),
        ("optimization_buffer_us", ctypes.c_double),
        ("source_copy_us", ctypes.c_double),
        ("binning_us", ctypes.c_double),
        ("march_us", ctypes.c_double),
        ("ssao_us", ctypes.c_double),
        ("shade_us", ctypes.c_double),
    ]


_lib = None

P = ctypes.c_void_p
I32 = ctypes.c_int32
F = ctypes.c_float

_SIGS = {
    "isaac_hip_create": (ctypes.c_int, [P, ctypes.c_int, P]),
    "isaac_hip_destroy": (None, [P]),
    "isaac_hip_last_error": (ctypes.c_char_p, [P]),
    "isaac_hip_set_stream": (ctypes.c_int, [P, P]),
    "isaac_hip_set_volume_source": (ctypes.c_int, [P, ctypes.c_int, P]),
    "isaac_hip_set_particle_source": (ctypes.c_int, [P, ctypes.c_int, P]),
    "isaac_hip_set_function": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_char_p, P]),
    "isaac_hip_set_transfer_points": (ctypes.c_int, [P, ctypes.c_int, ctypes.c_int, P, P]),
    "isaac_hip_set_transfer_lut": (ctypes.c_int, [P, ctypes.c_int, P]),
    "isaac_hip_set_weight": (ctypes.c_int, [P, ctypes.c_int, F]),
    "isaac_hip_set_iso_threshold": (ctypes.c_int, [P, ctypes.c_int, F]),
    "isaac_hip_set_step": (ctypes.c_int, [P, F]),
    "isaac_hip_set_interpolation": (ctypes.c_int, [P, ctypes.c_int]),
    "isaac_hip_set_dither_mode": (ctypes.c_int, [P, ctypes.c_int]),
    "isaac_hip_set_ao": (ctypes.c_int, [P, ctypes.c_int, F]),
    "isaac_hip_set_render_mode": (ctypes.c_int, [P, ctypes.c_int]),
    "isaac_hip_add_clipping": (ctypes.c_int, [P, F, F, F, F, F, F]),
    "isaac_hip_edit_clipping": (ctypes.c_int, [P, ctypes.c_uint32, F, F, F, F, F, F]),
    "isaac_hip_remove_clipping": (ctypes.c_int, [P, ctypes.c_uint32]),
    "isaac_hip_clipping_count": (ctypes.c_int, [P]),
    "isaac_hip_update_position": (ctypes.c_int, [P, P]),
    "isaac_hip_update_local_size": (ctypes.c_int, [P, P]),
    "isaac_hip_update_local_particle_size": (ctypes.c_int, [P, P]),
    "isaac_hip_update_buffers": (ctypes.c_int, [P]),
    "isaac_hip_render": (ctypes.c_int, [P, P, P, P, P, P, ctypes.c_int]),
    "isaac_hip_synchronize": (ctypes.c_int, [P]),
    "isaac_hip_min_max": (ctypes.c_int, [P, P, P, ctypes.c_int32]),
    "isaac_hip_set_advection": (ctypes.c_int, [P, ctypes.c_float, ctypes.c_float, ctypes.c_int32, ctypes.c_int32]),
    "isaac_hip_set_seed_points": (ctypes.c_int, [P, ctypes.c_int32]),
    "isaac_hip_get_advection": (ctypes.c_int, [P, ctypes.c_int, P, P, P]),
    "isaac_hip_get_timings": (ctypes.c_int, [P, P]),
    "isaac_hip_set_counting": (ctypes.c_int, [P, ctypes.c_int]),
    "isaac_hip_set_tuning": (ctypes.c_int, [P, I32, I32]),
    "isaac_hip_set_render_form": (ctypes.c_int, [P, I32, I32]),
    "isaac_hip_get_bricks": (ctypes.c_int, [P, ctypes.c_int, P, P, P, P]),
    "isaac_hip_set_float_output": (ctypes.c_int, [P, ctypes.c_int]),
    "isaac_hip_get_float_image": (ctypes.c_int, [P, P]),
    "isaac_hip_set_ao_mode": (ctypes.c_int, [P, I32]),
    "isaac_hip_batch_create": (ctypes.c_int, [P, I32, I32, I32, P]),
    "isaac_hip_batch_destroy": (None, [P]),
    "isaac_hip_batch_last_error": (ctypes.c_char_p, [P]),
    "isaac_hip_batch_set_tuning": (ctypes.c_int, [P, I32, I32]),
    "isaac_hip_batch_set_views": (ctypes.c_int, [P, I32, P, P, P]),
    "isaac_hip_batch_render": (ctypes.c_int, [P, I32]),
    "isaac_hip_batch_synchronize": (ctypes.c_int, [P]),
    "isaac_hip_batch_last_time_us": (ctypes.c_int, [P, P]),
    "isaac_hip_batch_get_images": (ctypes.c_int, [P, P, P]),
    "isaac_hip_composite_order": (ctypes.c_int, [ctypes.c_int32, P, P, P, P, P, P]),
    "isaac_hip_composite": (ctypes.c_int, [ctypes.c_int32, P, P, ctypes.c_int64, P, P]),
    "isaac_hip_base64": (ctypes.c_int, [P, ctypes.c_int64, P, P]),
    "isaac_hip_frame_message": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, P, P, ctypes.c_int64, P, P]),
    "isaac_hip_jpeg_workspace_bytes": (ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    "isaac_hip_video_frame_message": (ctypes.c_int, [P, ctypes.c_int32, P, ctypes.c_int64, P]),
    "isaac_hip_video_frame_capacity": (ctypes.c_int64, [P]),
    "isaac_hip_jpeg_encode": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_int64, P,
                                             ctypes.c_int64, P, P]),
    "isaac_hip_jpeg_frame_message": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_int64, P,
                                                    ctypes.c_int64, P, P]),
    "isaac_hip_jpeg_encode_noalloc": (ctypes.c_int, [P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, P, ctypes.c_int64, P,
                                                     P]),
    "isaac_hip_aonet_param_count": (ctypes.c_int64, []),
    "isaac_hip_aonet_set_params": (ctypes.c_int, [P, P, ctypes.c_int64]),
    "isaac_hip_aonet_get_params": (ctypes.c_int, [P, P, ctypes.c_int64]),
    "isaac_hip_aonet_clear_batch": (ctypes.c_int, [P]),
    "isaac_hip_aonet_add_crops": (ctypes.c_int, [P, I32, P, I32]),
    "isaac_hip_aonet_train_step": (ctypes.c_int, [P, F, P]),
    "isaac_hip_aonet_loss_grad": (ctypes.c_int, [P, I32, I32, I32, P, P, P, P]),
    "isaac_hip_aonet_forward": (ctypes.c_int, [P, I32, I32, I32, P, P]),
    "isaac_hip_get_counters": (ctypes.c_int, [P, P]),
    "isaac_hip_get_gbuffer": (ctypes.c_int, [P, P, P, P, P]),
    "isaac_hip_get_combined_textures": (ctypes.c_int, [P, P, P]),
    "isaac_hip_get_cell_list": (ctypes.c_int, [P, ctypes.c_int, P, P, P]),
    "isaac_hip_default_projection": (None, [I32, I32, P]),
    "isaac_hip_modelview": (None, [P, P, ctypes.c_double, P]),
    "isaac_hip_parse_functor_chain": (I32, [ctypes.c_char_p, P, P]),
    "isaac_hip_build_transfer_lut": (ctypes.c_int, [I32, I32, P, P, P]),
    "isaac_hip_default_transfer": (ctypes.c_int, [I32, I32, I32, I32, P, P]),
    "isaac_hip_version": (ctypes.c_char_p, []),
    "isaac_hip_copy_to_host": (ctypes.c_int, [P, P, P, ctypes.c_size_t]),
    "isaac_hip_get_clipping": (ctypes.c_int, [P, ctypes.c_uint32, P, P]),
    "isaac_hip_global_size_scaled": (ctypes.c_int, [P, P]),
    "isaac_hip_session_create": (ctypes.c_int, [P]),
    "isaac_hip_session_destroy": (None, [P]),
    "isaac_hip_register_message": (ctypes.c_int, [P, P, ctypes.c_char_p, I32, P, P, P, ctypes.c_int64, P]),
    "isaac_hip_period_message": (ctypes.c_int, [P, P, ctypes.c_int64, ctypes.c_uint32, P, ctypes.c_int64, P]),
    "isaac_hip_exit_message": (ctypes.c_int, [P, ctypes.c_int64, P]),
    "isaac_hip_server_send_bytes": (ctypes.c_int, [I32, ctypes.c_char_p, I32, P, ctypes.c_int64, P]),
}


class SteeringState(ctypes.Structure):
    """isaac_hip_steering_state (include/isaac_hip.h)."""

    _fields_ = [
        ("n_sources", ctypes.c_int32),
        ("n_volume_field", ctypes.c_int32),
        ("source_names", ctypes.POINTER(ctypes.c_char_p)),
        ("feature_dims", ctypes.POINTER(ctypes.c_int32)),
        ("functions", ctypes.POINTER(ctypes.c_char_p)),
        ("function_errors", ctypes.POINTER(ctypes.c_int32)),
        ("transfer_size", ctypes.c_int32),
        ("transfer_lut", ctypes.c_void_p),
        ("transfer_point_counts", ctypes.POINTER(ctypes.c_int32)),
        ("transfer_point_values", ctypes.POINTER(ctypes.c_uint32)),
        ("transfer_point_rgba", ctypes.POINTER(ctypes.c_float)),
        ("weights", ctypes.POINTER(ctypes.c_float)),
        ("iso_thresholds", ctypes.POINTER(ctypes.c_float)),
        ("min_max", ctypes.POINTER(ctypes.c_float)),
        ("step", ctypes.c_float),
        ("interpolation", ctypes.c_int32),
        ("render_optimization", ctypes.c_int32),
        ("seed_points", ctypes.c_int32),
        ("advection_step", ctypes.c_float),
        ("advection_weight", ctypes.c_float),
        ("advection_seeding_period", ctypes.c_int32),
        ("advection_seeding_duration", ctypes.c_int32),
        ("advection_on_pause", ctypes.c_int32),
        ("advection_border", ctypes.c_int32),
        ("background", ctypes.c_float * 3),
        ("n_clipping", ctypes.c_int32),
        ("clipping_position", ctypes.POINTER(ctypes.c_float)),
        ("clipping_normal", ctypes.POINTER(ctypes.c_float)),
        ("ao_enabled", ctypes.c_int32),
        ("ao_weight", ctypes.c_float),
        ("render_mode", ctypes.c_int32),
        ("dither_mode", ctypes.c_int32),
        ("projection", ctypes.c_double * 16),
        ("look_at", ctypes.c_double * 3),
        ("rotation", ctypes.c_double * 9),
        ("distance", ctypes.c_double),
        ("metadata", ctypes.c_char_p),
    ]


def header_functions() -> List[str]:
    """Names of every function declared in include/isaac_hip.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(isaac_hip_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IsaacHipError(
            f"{LIB_PATH} is missing: build the gfx950 library first (python -m isaac_amd.build); "
            "there is no fallback renderer"
        )
    # PyTorch-ROCm wheels carry their own libamdhip64/libhsa-runtime64.  Loading
    # torch first makes this library's HIP runtime bind to the HSA runtime that
    # torch already loaded (same soname libhsa-runtime64.so.1), so the process
    # has one HSA runtime and torch's device buffers are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    l = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(l, name)
        fn.restype = res
        fn.argtypes = args
    _lib = l
    return l


def check(rc: int, ctx=None, what: str = "") -> int:
    if rc < 0:
        msg = ""
        if ctx:
            m = lib().isaac_hip_last_error(ctx)
            msg = m.decode() if m else ""
        raise IsaacHipError(f"{what} failed: {ERRORS.get(rc, rc)} {msg}".strip())
    return rc
