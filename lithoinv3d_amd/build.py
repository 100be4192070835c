"""Builds liblithofwd.so (gfx950) in-tree with hipcc.

    python -m lithoinv3d_amd.build          # or build() from __graft_entry__

All HIP code is compiled with packed-FP32 instruction selection disabled
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblithofwd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # code object v5: loadable by this image's ROCm 7.2 runtime and by the ROCm 7.0
    # runtime bundled with PyTorch (one runtime per process, see lithofwd.load_library)
    "-mcode-object-version=5",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
    "-Wno-unused-value",
    "-Wno-unused-result",
]


HOST = os.path.join(HERE, "host")
HOST_LIB = os.path.join(HERE, "liblithohost.so")
CXX = os.environ.get("CXX", "g++")


def sources():
    return [os.path.join(CSRC, "lithofwd.hip"), os.path.join(CSRC, "mlp.hip")]


def build_host(verbose: bool = False) -> str:
    """The C++ host layer (ForwardModel / objectives / L-BFGS-B) over the C ABI,
    linked against liblithofwd.so."""
    srcs = [os.path.join(HOST, f) for f in ("litho_host.cpp", "litho_io.cpp", "litho_runner.cpp",
                                            "litho_capi.cpp")]
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-Wextra",
           "-Wno-unused-parameter"] + srcs + [
        "-L" + HERE, "-llithofwd", "-Wl,-rpath,$ORIGIN", "-o", HOST_LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=HOST)
    os.replace(HOST_LIB + ".tmp", HOST_LIB)
    return HOST_LIB


def build(verbose: bool = False, extra=None, out: str = LIB) -> str:
    cmd = [HIPCC] + HIP_FLAGS + list(extra or []) + sources() + ["-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=CSRC)
    os.replace(out + ".tmp", out)
    if out == LIB:
        build_host(verbose)
    return out


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
    print("built", LIB)
