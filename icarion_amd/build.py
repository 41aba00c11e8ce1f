"""Build recipe of the gfx950 library: hipcc -> icarion_amd/lib/libicarion_gpu.so.

Runs on a machine without a GPU (hipcc cross-compiles for gfx950).  Every HIP
translation unit is compiled with packed-FP32 instruction selection disabled.
The translation units share no device code across files (the device building
blocks are headers), so each is compiled to its own object in parallel and the
objects are linked into one shared library.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(LIBDIR, "obj")
LIB = os.path.join(LIBDIR, "libicarion_gpu.so")
SOURCES = ["step_kernels.hip", "rk45_kernels.hip", "multi_kernels.hip", "metrics_kernels.hip",
           "ehss_tools.hip", "surrogate.hip", "order_kernels.hip", "icarion_gpu_api.cpp"]
ARCH = os.environ.get("ICARION_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (set HIPCC)")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"{os.path.basename(cmd[0])} failed with exit code {res.returncode}: {cmd[-1]}")
    return res


# The trajectory file writer (csrc/trajectory_h5.cpp, include/icarion_h5.h) is a
# host library of its own, linked against the HDF5 C library; the GPU library
# does not depend on it.  HDF5_DIR (default /opt/conda) holds include/hdf5.h
# and lib/libhdf5.so.
H5_LIB = os.path.join(LIBDIR, "libicarion_h5.so")


def hdf5_dir():
    d = os.environ.get("HDF5_DIR", "/opt/conda")
    if os.path.exists(os.path.join(d, "include", "hdf5.h")) and os.path.exists(
            os.path.join(d, "lib", "libhdf5.so")):
        return d
    return None


def build_h5(verbose=False):
    """Build libicarion_h5.so; returns its path, or None without an HDF5 library."""
    d = hdf5_dir()
    if d is None:
        if verbose:
            print("HDF5 not found: libicarion_h5.so not built", flush=True)
        return None
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = H5_LIB + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
          "-I", os.path.join(d, "include"), os.path.join(CSRC, "trajectory_h5.cpp"),
          "-L", os.path.join(d, "lib"), "-lhdf5", f"-Wl,-rpath,{os.path.join(d, 'lib')}",
          "-o", tmp], verbose)
    os.replace(tmp, H5_LIB)
    return H5_LIB


def build(verbose=False, extra_flags=(), jobs=None):
    os.makedirs(OBJDIR, exist_ok=True)
    cc = hipcc()
    sources = list(SOURCES)

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        cmd = [cc, f"--offload-arch={ARCH}"] + HIP_FLAGS + list(extra_flags) + [
            "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        _run(cmd, verbose)
        return obj

    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sources))
    tmp = LIB + ".tmp"
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + objs + ["-o", tmp], verbose)
    os.replace(tmp, LIB)
    build_h5(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
