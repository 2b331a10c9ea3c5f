"""The march kernels replace IEEE x/255 by a reciprocal product with one fma
correction (isaac_amd/csrc/render.hip, div255).  This checks the identity for
every binary32 input in [0, 256] (trilinear results of RGBA8 texels lie in
[0, 255]), compiled with -ffp-contract=off like the library."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_div255_identity_exhaustive(tmp_path):
    exe = tmp_path / "div255_check"
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), os.path.join(HERE, "cpp", "div255_check.c"), "-lm"])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout
