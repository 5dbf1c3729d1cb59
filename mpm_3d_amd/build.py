"""Build the native libraries in-tree.

* ``mpm_3d_amd/lib/libmpm_gpu.so`` -- the HIP kernels and the C-ABI (include/mpm_gpu.h),
  compiled for gfx950 only.
* ``mpm_3d_amd/lib/libmpm_host.so`` and ``mpm_3d_amd/lib/mpm_v3_hip`` -- the C++ host
  framework (config grammar, registry, job, plugins, HipExplicitUSL solver) linked
  against libmpm_gpu.so.

Run ``python -m mpm_3d_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "mpm_3d_amd")
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
LIB = os.path.join(PKG, "lib")
INC = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")  # roctx ranges (rocprofiler-sdk-roctx)
ARCH = "gfx950"

# -ffp-contract=off: no implicit multiply-add fusion, so the connectivity arithmetic
# (corner positions, element indices) rounds exactly like the reference; hot loops use
# explicit fma(). Packed FP32 instructions are disabled for all device code.
# The atomic optimizer's DPP strategy reduces same-address atomics (the LDS bounding-box
# min/max of a particle block) with a few DPP steps; the default iterative strategy
# loops over every active lane on the scalar unit, thousands of SALU ops per point.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-munsafe-fp-atomics", "-mllvm", "-amdgpu-atomic-optimizer-strategy=DPP",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_gpu(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmpm_gpu.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, f"-I{INC}", f"-I{CSRC}", "-o", tmp, os.path.join(CSRC, "mpm_gpu.hip"),
          f"-L{ROCM_LIB}", "-lrocprofiler-sdk-roctx", f"-Wl,-rpath,{ROCM_LIB}"])
    os.replace(tmp, out)
    return out


def build_host():
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cpp")))
    if not srcs:
        return None
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libmpm_host.so")
    common = ["g++", "-O2", "-std=c++17", "-fPIC", f"-I{INC}", f"-I{HOST}"]
    lib_srcs = [s for s in srcs if not s.endswith("main.cpp")]
    _run([*common, "-shared", "-o", out + ".tmp", *lib_srcs, f"-L{LIB}", "-lmpm_gpu",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    mains = [s for s in srcs if s.endswith("main.cpp")]
    if mains:
        exe = os.path.join(LIB, "mpm_v3_hip")
        _run([*common, "-o", exe + ".tmp", *mains, f"-L{LIB}", "-lmpm_host", "-lmpm_gpu",
              "-Wl,-rpath,$ORIGIN"])
        os.replace(exe + ".tmp", exe)
    return out


def build_all():
    build_gpu()
    build_host()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
