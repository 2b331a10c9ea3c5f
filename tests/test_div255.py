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


def test_general_division_identity_sampled(tmp_path):
    """The classification kernel divides by per-source constants (iso thresholds,
    the sum of source weights), and the brick build by each brick's scale, with the
    same one-correction scheme and r = RN(1/d) (the sign of x restored for -0):
    2e8 seeded (x, d) pairs, d over [1e-3, 1e3] and extreme mantissas, x of both
    signs and +-0, a quarter of them next to rounding boundaries, all equal to IEEE
    x / d bit for bit."""
    exe = tmp_path / "divq_check"
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), os.path.join(HERE, "cpp", "divq_check.c"), "-lm"])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0" in r.stdout


def test_rounding_shortcuts_exhaustive(tmp_path):
    """Classification and the transfer-function lookups replace roundf by
    trunc(x + 0x1.fffffep-2) (byte quantisation of [0, 1] x 255; with 0.5 in place of
    the largest binary32 below it the check must find differences) and
    trunc(med3(x + 0x1.fffffep-2, 0, T - 1)) (transfer index): every binary32
    input in the ranges where the two forms could differ."""
    exe = tmp_path / "round_check"
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), os.path.join(HERE, "cpp", "round_check.c"), "-lm"])
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "mismatches 0;" in r.stdout and r.stdout.rstrip().endswith("mismatches 0")
