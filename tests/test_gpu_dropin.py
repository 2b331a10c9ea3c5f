"""The C++ drop-in (include/isaac_mi355x.hpp) end to end on the GPU: the
example application renders the small workload through IsaacVisualization and
its framebuffer is compared with the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from isaac_amd import scenes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_dropin_example_matches_oracle(tmp_path, oracle_lib, hip_lib):
    from isaac_amd import build

    exe = build.SHIM_BIN
    if not os.path.exists(exe):
        build.build_shim_example()
    spec = scenes.small_scene()
    field = tmp_path / "field.raw"
    out = tmp_path / "frame.rgba"
    spec.volumes[0].data.astype(np.float32).tofile(field)
    W, H = spec.framebuffer
    print("running", exe, flush=True)
    r = subprocess.run([exe, str(field), "64", str(W), str(H), str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    img = np.fromfile(out, np.uint8).reshape(H, W, 4)
    ora = oracle_lib.render(spec, 32)
    d = np.abs(img.astype(int) - ora.rgba.astype(int)).max(axis=2)
    print(f"c++ drop-in: {(d <= 2).mean() * 100:.4f}% within 2/255, max {d.max()}")
    assert (d <= 2).mean() >= 0.999
    # the min/max request through the drop-in (sendMinMax -> minMaxArray)
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("minmax ")][0].split()
    mins, maxs = oracle_lib.min_max(spec)
    assert np.float32(line[1]) == mins[0] and np.float32(line[2]) == maxs[0], (line, mins, maxs)
    # the video frame message (videoFrameMessage -> the JPEG path of serverSendFrame) of that frame
    from oracle import jpeg_oracle

    msg = (tmp_path / "frame.rgba.msg").read_bytes()
    assert msg == jpeg_oracle.frame_message(img, 90)
    # the server messages of the drop-in against the message oracle, from the state the
    # example set: one density source, its 4-point transfer function, interpolation,
    # step 0.5, black background, the default camera and the min/max it requested
    from oracle import message_oracle as M

    pts = [(0, (0.0, 0.0, 0.4, 0.0)), (256, (0.0, 0.5, 1.0, 0.05)), (614, (1.0, 0.85, 0.1, 0.35)),
           (1024, (1.0, 0.15, 0.1, 0.9))]
    st = M.State(
        source_names=["density"], feature_dims=[1], n_volume_field=1, functions=["idem"], function_errors=[0],
        transfer_size=1024, transfer_lut=oracle_lib.build_lut(pts, 1024)[None], transfer_points=[pts],
        weights=[2.0], iso_thresholds=[0.0], min_max=[(float(line[1]), float(line[2]))], step=0.5, interpolation=True,
        background=(0.0, 0.0, 0.0), projection=[float(x) for x in oracle_lib.default_projection(W, H)],
    )
    sess = M.Session()
    assert (tmp_path / "frame.rgba.register").read_text() == sess.register(st, "shim_example", 1, (W, H), (64, 64, 64))
    assert (tmp_path / "frame.rgba.period").read_text() == sess.period(st, 0, M.SEND)
    assert (tmp_path / "frame.rgba.exit").read_text() == M.exit_message()
