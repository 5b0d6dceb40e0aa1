"""CPU sanitizer builds of the host-only C++: the library's parsers, generators,
exp(K) and reverse-communication Lanczos (ciflow_amd/csrc/host.cpp,
integrals.cpp, lanczos.cpp) and the oracle (oracle/ciflow_oracle.cpp), and the
HDF5 file library (ciflow_amd/csrc/h5io.cpp, linked against the HDF5 C
library), compiled with g++ under AddressSanitizer + UndefinedBehaviorSanitizer
(with leak detection) and run over the checks in tests/sanitize/.  Code that
needs the GPU is not part of these builds; the GPU tests cover it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ciflow_amd", "csrc")
SOURCES = [os.path.join(ROOT, "tests", "sanitize", "host_checks.cpp"), os.path.join(CSRC, "host.cpp"),
           os.path.join(CSRC, "integrals.cpp"), os.path.join(CSRC, "lanczos.cpp"),
           os.path.join(ROOT, "oracle", "ciflow_oracle.cpp")]
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fopenmp", "-ffp-contract=off",
         "-fno-builtin-sin", "-fno-builtin-cos", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
         "-Wno-unknown-pragmas"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_address_and_undefined_behavior_sanitizers(tmp_path):
    exe = tmp_path / "host_checks"
    r = subprocess.run(["g++", *FLAGS, "-o", str(exe), *SOURCES], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "all host checks passed" in r.stdout, out[-6000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-6000:]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_hdf5_file_library_under_address_and_undefined_behavior_sanitizers(tmp_path):
    from ciflow_amd.build import hdf5_link_args, hdf5_prefix

    prefix = hdf5_prefix()
    if prefix is None:
        pytest.skip("no HDF5 C library")
    exe = tmp_path / "h5_checks"
    src = [os.path.join(ROOT, "tests", "sanitize", "h5_checks.cpp"), os.path.join(CSRC, "h5io.cpp")]
    r = subprocess.run(["g++", *FLAGS, "-I", os.path.join(prefix, "include"), "-o", str(exe), *src,
                        *hdf5_link_args(prefix)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([str(exe), os.path.join(ROOT, "tests", "golden"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0 and "all h5 checks passed" in r.stdout, out[-6000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-6000:]
