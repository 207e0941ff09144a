"""Build recipe of the HIP library (libucmc.so, gfx950) and the C++ host-layer
programs.  Called by __graft_entry__.build(); always recompiles."""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libucmc.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: packed FP32 forms are kept out of every kernel.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++20", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include"), "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return r


def build_library(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    objs = [os.path.join(OBJ, os.path.basename(s) + ".o") for s in srcs]
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda so: _run([HIPCC, *HIP_FLAGS, "-c", so[0], "-o", so[1]]),
                    zip(srcs, objs)))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_host_programs():
    """C++23 test programs over the host layer (include/ucmc/gpu.hpp), linked
    against libucmc.so; host code only, compiled with g++."""
    os.makedirs(OBJ, exist_ok=True)
    out = os.path.join(OBJ, "host_layer_check")
    _run(["g++", "-std=c++23", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
          "-I" + os.path.join(HERE, "include"), os.path.join(ROOT, "tests", "cpp", "host_layer_check.cpp"),
          "-L" + HERE, "-lucmc", "-Wl,-rpath,$ORIGIN/..", "-o", out])
    return [out]


if __name__ == "__main__":
    print(build_library())
    print(build_host_programs())
