"""The C-ABI library on the CPU: it loads, exports every function that
include/isaac_hip.h declares, validates configurations before touching a
device, and the C++ drop-in header builds against it."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_function(hip_lib):
    from isaac_amd import _lib

    declared = _lib.header_functions()
    assert len(declared) >= 39
    missing = [f for f in declared if not hasattr(hip_lib, f)]
    assert not missing, missing
    assert set(_lib._SIGS) == set(declared), set(_lib._SIGS) ^ set(declared)
    assert hip_lib.isaac_hip_version().startswith(b"isaac-hip")


def _cfg(**kw):
    from isaac_amd import _lib

    c = _lib.Config()
    c.framebuffer_size[:] = (64, 48)
    c.global_size[:] = (16, 16, 16)
    c.local_size[:] = (16, 16, 16)
    c.local_particle_size[:] = (4, 4, 4)
    c.position[:] = (0, 0, 0)
    c.scale[:] = (1.0, 1.0, 1.0)
    c.transfer_size = 1024
    c.n_volume_sources = 1
    c.n_field_sources = 0
    c.n_particle_sources = 0
    for k, v in kw.items():
        if isinstance(v, tuple):
            getattr(c, k)[:] = v
        else:
            setattr(c, k, v)
    return c


@pytest.mark.parametrize(
    "kw,rc",
    [
        (dict(n_field_sources=-1), -1),
        (dict(n_volume_sources=10, n_field_sources=7), -1),  # volume + field sources above 16
        (dict(transfer_size=0), -1),
        (dict(framebuffer_size=(0, 10)), -1),
        (dict(local_size=(0, 16, 16)), -1),
        (dict(scale=(1.0, -1.0, 1.0)), -1),
        (dict(n_volume_sources=17), -1),
    ],
)
def test_create_rejects_bad_configs_before_device(hip_lib, kw, rc):
    ctx = ctypes.c_void_p()
    cfg = _cfg(**kw)
    assert hip_lib.isaac_hip_create(ctypes.byref(cfg), 0, ctypes.byref(ctx)) == rc
    assert not ctx.value


def test_null_arguments(hip_lib):
    assert hip_lib.isaac_hip_create(None, 0, None) == -1
    assert hip_lib.isaac_hip_set_step(None, 0.5) == -1
    assert hip_lib.isaac_hip_render(None, None, None, None, None, None, 0) == -1
    assert hip_lib.isaac_hip_last_error(None) == b"null context"
    assert hip_lib.isaac_hip_build_transfer_lut(0, 0, None, None, None) == -1


def test_cpp_dropin_header_builds(hip_lib):
    from isaac_amd import build

    path = build.build_shim_example()
    assert path and os.path.exists(path)


def test_cpp_dropin_header_is_hip_free():
    """The drop-in header includes only the C-ABI: no HIP, torch or glm types."""
    text = open(os.path.join(ROOT, "include", "isaac_mi355x.hpp")).read()
    for bad in ("hip_runtime", "torch", "glm/", "alpaka"):
        assert f"#include <{bad}" not in text and f'#include "{bad}' not in text
    text = open(os.path.join(ROOT, "include", "isaac_hip.h")).read()
    assert "#include <hip" not in text


def test_python_constants_match_the_header():
    """The Python mirror's tuning keys, render forms and occlusion modes carry the
    values include/isaac_hip.h gives them."""
    import re

    from isaac_amd.visualization import IsaacVisualization as V

    text = open(os.path.join(ROOT, "include", "isaac_hip.h")).read()
    found = 0
    for prefix in ("TUNE", "FORM", "AO"):
        for name, value in re.findall(rf"ISAAC_HIP_({prefix}_\w+)\s*=\s*(\d+)", text):
            assert getattr(V, name) == int(value), name
            found += 1
    assert found >= 10
