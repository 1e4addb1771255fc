"""Build the in-tree native libraries.

  lib/libhfv.so       hipcc, gfx950: device kernels + C-ABI (include/hfv.h)
  lib/libhfv_host.so  g++ + OpenMP: host tessellation (include/hfv_host.h)

Usage: python -m hydro_rs_amd.build   (or __graft_entry__.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# FP64 parity needs the reference's operation order: no FMA contraction.  Packed FP32 instructions are
# disabled for the whole device build.
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd)


def build_host():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libhfv_host.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall", "-o", out,
          os.path.join(CSRC, "host", "voronoi.cpp")])
    return out


def build_hip():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libhfv.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-o", tmp, os.path.join(CSRC, "hfv_api.hip")])
    os.replace(tmp, out)
    return out


def build_all():
    build_host()
    build_hip()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"host": build_host, "hip": build_hip, "all": build_all}[what]()
