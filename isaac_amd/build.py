"""Builds the gfx950 HIP library isaac_amd/libisaac_hip.so in-tree.

    python -m isaac_amd.build          # or isaac_amd.build.build()

Flags:
  --offload-arch=gfx950                   MI355X (CDNA4) only
  -ffp-contract=off                       no FMA contraction: every binary32 + - * / is one
                                          IEEE op in the reference's order (parity path)
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                          no packed FP32 (v_pk_*_f32) instructions
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libisaac_hip.so")
SOURCES = ["host.cpp", "aonet_host.cpp", "batch_host.cpp", "render.hip", "particles.hip", "aonet.hip", "composite.hip", "frame.hip", "jpeg.hip", "messages.cpp"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def flags() -> list:
    return [
        "-O3",
        "-std=c++17",
        "--offload-arch=gfx950",
        "-fPIC",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-result",
        "-I" + os.path.join(ROOT, "include"),
        "-I" + CSRC,
    ]


def build(verbose: bool = False) -> str:
    objs = []
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cc = hipcc()
    procs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [cc, *flags(), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = False
    for cmd, p in procs:
        out, _ = p.communicate()
        if out and (verbose or p.returncode != 0):
            sys.stdout.write(out.decode(errors="replace"))
        if p.returncode != 0:
            failed = True
    if failed:
        raise RuntimeError("HIP compilation failed")
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.check_call(cmd)
    os.replace(tmp, LIB)
    return LIB


SHIM_SRC = os.path.join(ROOT, "tests", "cpp", "shim_example.cpp")
SHIM_BIN = os.path.join(ROOT, "tests", "cpp", "shim_example")


def build_shim_example() -> str:
    """Compile the C++ IsaacVisualization shim example (include/isaac_mi355x.hpp)
    against the library with the host compiler only: the shim never includes
    HIP headers, it calls the C-ABI."""
    if not os.path.exists(SHIM_SRC):
        return ""
    cxx = shutil.which("g++") or "c++"
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(rocm, "include"), SHIM_SRC, "-o", SHIM_BIN,
           "-L" + HERE, "-lisaac_hip", "-Wl,-rpath," + HERE, "-L" + os.path.join(rocm, "lib"), "-lamdhip64"]
    subprocess.check_call(cmd)
    return SHIM_BIN


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
