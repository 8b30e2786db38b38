"""Build the HIP shared library in-tree for gfx950.

``python -m frontierlabproject_amd.build_native`` compiles
``csrc/*.hip`` and ``csrc/*.cpp`` with hipcc into ``lib/libqgpusim_hip.so``.
All device code is built with packed FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIBDIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIBDIR, "libqgpusim_hip.so")
ARCH = "gfx950"

SOURCES = ["qsim_kernels.hip", "qsim_capi.cpp", "qsim_plan.cpp"]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build the backend")


def build(verbose=False, extra_flags=(), out=None):
    out = out or LIB
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp%d" % os.getpid()
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           "-I", os.path.join(ROOT, "include"), "-I", CSRC, "-o", tmp]
    cmd += list(extra_flags)
    cmd += ["-x", "hip"] + [os.path.join(CSRC, s) for s in SOURCES]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building %s" % out)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose=True))
