"""GPU test of the C++ bindings a CIFlow maintainer would add (Routes 1, 2 and 3),
compiled with g++ against include/ciflow_amd.h by __graft_entry__.build()
(tests/integration/Makefile) and run as its own process: the RCI loop
driving a GPU mvprod through a virtual mvprod seam, cf_solve, cf_rdm and the
32-bit row-pointer export on FCI L=8 4a4b, checked against the dense-eigh
energy of the oracle's matrix."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "integration", "routes")


def test_cpp_routes_against_the_c_abi():
    assert os.path.exists(EXE), "tests/integration/routes is not built: run __graft_entry__.build()"
    print("running", EXE, flush=True)
    sys.stdout.flush()
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "routes ok" in r.stdout, r.stdout + r.stderr
