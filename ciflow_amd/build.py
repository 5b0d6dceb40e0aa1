"""Build libciflow_amd.so in-tree with hipcc for gfx950.

Usage: python -m ciflow_amd.build  (or __graft_entry__.build()).
Objects go to ciflow_amd/_build/, the shared library to ciflow_amd/.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libciflow_amd.so")
ARCH = os.environ.get("CIFLOW_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
          "-I", os.path.join(ROOT, "include"), "-I", CSRC]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

SOURCES = ["strings.hip", "build.hip", "spmv.hip", "spmv_fci.hip", "spmv_ell.hip", "spmv_doci.hip", "krylov.hip", "rdm.hip", "transform.hip", "sigma.hip", "noalloc.hip", "batch.hip", "davbatch.hip",
           "lanczos.cpp", "integrals.cpp", "host.cpp", "capi.cpp", "solve_batch.cpp", "server.cpp"]


def _compile(src: str) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC, "-x", "hip"] + COMMON + DEVICE + ["-c", path, "-o", obj]
    else:
        cmd = [HIPCC] + COMMON + [f"--offload-arch={ARCH}", "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    warn = "\n".join(ln for ln in r.stderr.splitlines() if "packed-fp32-ops" not in ln and ln.strip())
    if "warning" in warn:
        print(warn, file=sys.stderr)
    return obj


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = min(len(SOURCES), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-L/opt/rocm/lib", "-lrocblas",
                                                                                     "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


H5_LIB = os.path.join(HERE, "libciflow_h5.so")
H5_PREFIXES = [os.environ.get("CIFLOW_HDF5_PREFIX", ""), "/opt/conda", "/usr"]


def hdf5_prefix() -> str | None:
    """Install prefix of the HDF5 C library (headers and libhdf5.so), or None."""
    for p in H5_PREFIXES:
        if p and os.path.exists(os.path.join(p, "include", "hdf5.h")) and \
                any(os.path.exists(os.path.join(p, d, "libhdf5.so")) for d in ("lib", "lib64")):
            return p
    return None


def hdf5_link_args(prefix: str) -> list[str]:
    """Linker arguments for the HDF5 C library under `prefix`: the library by
    its path (no -L, so the compiler's own C++ runtime is the one linked) and a
    run path that finds the system C++ runtime before anything else in the
    HDF5 prefix (an HDF5 distribution may ship an older libstdc++)."""
    libdir = os.path.join(prefix, "lib" if os.path.exists(os.path.join(prefix, "lib", "libhdf5.so")) else "lib64")
    gxx = shutil.which("g++") or "g++"
    r = subprocess.run([gxx, "-print-file-name=libstdc++.so"], capture_output=True, text=True)
    cxx_dir = os.path.dirname(os.path.realpath(r.stdout.strip())) if r.returncode == 0 else ""
    rpath = ":".join(d for d in (cxx_dir, libdir) if d)
    return [os.path.join(libdir, "libhdf5.so"), f"-Wl,-rpath,{rpath}"]


def build_h5(verbose: bool = True) -> str | None:
    """Build the host-only HDF5 file library libciflow_h5.so (no GPU code).
    Returns None, and builds nothing, where no HDF5 C library is installed."""
    prefix = hdf5_prefix()
    if prefix is None:
        if verbose:
            print("HDF5 C library not found: libciflow_h5.so not built")
        return None
    tmp = H5_LIB + ".tmp"
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I", os.path.join(prefix, "include"), os.path.join(CSRC, "h5io.cpp"), "-o", tmp] + hdf5_link_args(prefix)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"HDF5 library build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, H5_LIB)
    if verbose:
        print(f"built {H5_LIB}")
    return H5_LIB


if __name__ == "__main__":
    build()
    build_h5()
