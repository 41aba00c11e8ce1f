"""Build recipe of the gfx950 library: hipcc -> icarion_amd/lib/libicarion_gpu.so.

Runs on a machine without a GPU (hipcc cross-compiles for gfx950).  Every HIP
translation unit is compiled with packed-FP32 instruction selection disabled.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libicarion_gpu.so")
SOURCES = ["step_kernels.hip", "rk45_kernels.hip", "multi_kernels.hip", "metrics_kernels.hip", "ehss_tools.hip",
           "surrogate.hip", "icarion_gpu_api.cpp"]
ARCH = os.environ.get("ICARION_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def build(verbose=False, extra_flags=()):
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument",
           "-x", "hip"] + [os.path.join(CSRC, s) for s in SOURCES] + list(extra_flags) + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
