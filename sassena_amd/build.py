"""Build libsassena_amd.so in-tree with hipcc for gfx950.

    python -m sassena_amd.build          (or __graft_entry__.build())

The library has no torch dependency: it is the HIP kernels plus the C ABI of
include/sassena_amd.h.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB = os.path.join(PKG, "libsassena_amd.so")
ARCH = os.environ.get("SX_OFFLOAD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Wall",
    # keep hand-written FP32 code off the packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


HOST_LIB = os.path.join(PKG, "libsassena_host.so")
SX_RUN = os.path.join(PKG, "bin", "sx_run")
CONDA = os.environ.get("SX_HDF5_PREFIX", "/opt/conda")


def sources():
    return sorted(glob.glob(os.path.join(PKG, "csrc", "*.hip")))


def host_sources():
    return sorted(p for p in glob.glob(os.path.join(PKG, "host", "*.cpp")) if not p.endswith("sx_run.cpp"))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"{cmd[0]} failed ({res.returncode}):\n{res.stderr}")


def build_host(verbose: bool = False) -> str:
    """The C++ host layer (device classes, config, I/O) and the sx_run driver,
    linked against libsassena_amd.so and HDF5 (+HL)."""
    cxx = os.environ.get("CXX", "g++")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(PKG, "host"), "-I", os.path.join(CONDA, "include")]
    common = ["-O2", "-std=c++17", "-Wall", "-Wextra", "-fPIC"]
    # HDF5 by full path: a -L on the HDF5 prefix would let the linker pick
    # that prefix's older C++ runtime instead of the system one ROCm needs.
    h5 = [os.path.join(CONDA, "lib", "libhdf5_hl.so"), os.path.join(CONDA, "lib", "libhdf5.so")]
    libs = ["-L", PKG, "-lsassena_amd", *h5, "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{os.path.join(CONDA, 'lib')}"]
    tmp = HOST_LIB + ".tmp"
    _run([cxx, *common, "-shared", *inc, "-o", tmp, *host_sources(), *libs], verbose)
    os.replace(tmp, HOST_LIB)
    os.makedirs(os.path.dirname(SX_RUN), exist_ok=True)
    tmp = SX_RUN + ".tmp"
    _run([cxx, *common, *inc, "-o", tmp, os.path.join(PKG, "host", "sx_run.cpp"), "-L", PKG, "-lsassena_host",
          "-lsassena_amd", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath-link,{os.path.join(CONDA, 'lib')}"], verbose)
    os.replace(tmp, SX_RUN)
    return SX_RUN


def _hipcc(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    errs = "\n".join(l for l in (res.stderr or "").splitlines() if noise not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{errs}")
    if verbose and errs.strip():
        print(errs, file=sys.stderr)


def build(verbose: bool = False) -> str:
    """Every .hip source compiled to an object in parallel (each holds its own
    kernels and launchers; device code does not cross files), then linked."""
    from concurrent.futures import ThreadPoolExecutor
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    compile_flags = [f for f in HIPCC_FLAGS if f != "-shared"]
    jobs = []
    for src in sources():
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        jobs.append((obj, [hipcc, *compile_flags, "-c", "-I", os.path.join(ROOT, "include"), "-o", obj, src]))
    workers = max(1, min(len(jobs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(workers) as ex:
        list(ex.map(lambda j: _hipcc(j[1], verbose), jobs))
    tmp = LIB + ".tmp"
    _hipcc([hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in jobs]], verbose)
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
