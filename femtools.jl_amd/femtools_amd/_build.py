"""Compiles libfemdr.so (the HIP/gfx950 library behind include/femdr.h) in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)                 # femtools.jl_amd/
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
LIB = os.path.join(PKG, "libfemdr.so")
SOURCES = [os.path.join(CSRC, "femdr.hip"), os.path.join(CSRC, "femdr_thermal.hip"),
           os.path.join(CSRC, "femdr_hex27.hip"), os.path.join(CSRC, "femdr_q9.hip"), os.path.join(CSRC, "femcnn.hip")]
DEPS = SOURCES + [os.path.join(CSRC, "femdr_kernels.hpp"), os.path.join(CSRC, "femdr_host.hpp"),
                  os.path.join(REPO, "include", "femdr.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result", "-Wno-unused-value", "-Wno-unused-function", "-Wno-unused-variable",
]


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    cmd = [HIPCC, *FLAGS, *(f"-D{d}" for d in defines), "-I", os.path.join(REPO, "include"), *SOURCES,
           "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfemdr.so")
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
