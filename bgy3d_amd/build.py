"""Build the gfx950 shared library ``bgy3d_amd/librism3d.so`` in-tree.

    python -m bgy3d_amd.build

hipcc compiles the HIP translation unit for gfx950 only (cross-compiles
without a GPU); the host C++ of the 1D chi producer is built by hipcc as
plain C++.  Packed-FP32 instruction selection is disabled for device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librism3d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def build(verbose=False):
    objs = []
    srcs = [("solver.hip", True), ("rism1d.cpp", False)]
    for src, is_hip in srcs:
        out = os.path.join(CSRC, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC] + (HIP_FLAGS if is_hip else ["-O3", "-std=c++17", "-fPIC", "-Wall"])
        cmd += ["-c", os.path.join(CSRC, src), "-o", out]
        if not is_hip:
            cmd = [HIPCC, "-x", "c++"] + cmd[1:]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(out)
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", LIB + ".tmp"] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
