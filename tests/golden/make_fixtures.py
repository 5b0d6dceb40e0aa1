"""Convert the reference's psi4 integral dumps into compact .npz fixtures.

Reads the dump format the reference parses in Hamiltonian::read_file
(Hamiltonian.cpp:707-885): header lines Nalpha/Nbeta/Symmetry/Nirreps/Nuclear
Repulsion/Number Of Molecular Orbitals/Irreps, then "****  MO OEI" lines
"i j value" and "****  MO TEI" lines "i j k l value" in chemist notation, which
the reference stores as physics setVmat(i,k,j,l) (:873) in an 8-fold symmetric
container (FourIndex.h:25-31).  The fixtures hold only data: dense h[L][L],
dense physics V[L][L][L][L] as getVmat(i,j,k,l) returns it, econst, nalpha,
nbeta and the orbital irreps.

Run from the repo root:  python tests/golden/make_fixtures.py
(needs /root/reference; the generated .npz files are committed).  The two
determinant lists of the reference's selected-CI tests, the BeH2 dump, the two
HDF5 integral files, the RHF test's overlap and the Hamiltonian test's two
orbital transformations are copied unchanged.
"""
from __future__ import annotations

import os
import shutil
import sys

import numpy as np

REF = "/root/reference/tests/data"
OUT = os.path.dirname(os.path.abspath(__file__))


def read_dump(path):
    with open(path) as f:
        lines = f.read().splitlines()
    i = next(k for k, ln in enumerate(lines) if "****  Molecular Integrals" in ln)
    val = lambda ln: ln.split("=", 1)[1].strip()  # noqa: E731
    na = int(val(lines[i + 1]))
    nb = int(val(lines[i + 2]))
    econst = float(val(lines[i + 5]))
    L = int(val(lines[i + 6]))
    irreps = [int(x) for x in lines[i + 8].split()]
    # MO dumps mark the sections "****  MO OEI" / "****  MO TEI", the
    # atomic-orbital dump "*** OEI SO" / "*** TEI"
    k = next(k for k in range(i, len(lines)) if lines[k].startswith(("****  MO OEI", "*** OEI")))
    h = np.zeros((L, L))
    k += 1
    while not lines[k].startswith("*"):
        p = lines[k].split()
        a, b, v = int(p[0]), int(p[1]), float(p[2])
        h[a, b] = v
        h[b, a] = v
        k += 1
    assert lines[k].startswith(("****  MO TEI", "*** TEI"))
    k += 1
    V = np.zeros((L, L, L, L))
    while not lines[k].startswith("*"):
        p = lines[k].split()
        a, b, c, d, v = int(p[0]), int(p[1]), int(p[2]), int(p[3]), float(p[4])
        # chemist (ab|cd): all 8 chemist permutations, stored as physics <ac|bd>
        for (p0, q0, r0, s0) in ((a, b, c, d), (b, a, c, d), (a, b, d, c), (b, a, d, c),
                                 (c, d, a, b), (d, c, a, b), (c, d, b, a), (d, c, b, a)):
            V[p0, r0, q0, s0] = v
        k += 1
    return dict(L=L, nalpha=na, nbeta=nb, econst=econst, irreps=np.array(irreps), h=h, V=V)


def main():
    if not os.path.isdir(REF):
        print("reference data not found", file=sys.stderr)
        return 1
    for name, src in (("beh2_sto3g", "beh2_sto3g_1_5.dat"), ("psi0_631g", "psi0_6-31g0.60.mout"),
                      ("be_ccpvdz", "Be_ccpVDZ.dat"), ("beh2_atomic", "beh2atomic.dat")):
        d = read_dump(os.path.join(REF, src))
        np.savez_compressed(os.path.join(OUT, f"{name}.npz"), **d)
        print(name, d["L"], d["nalpha"], d["nbeta"], d["econst"])
    # determinant lists the reference's selected-CI tests read (Permutator_File)
    # and the BeH2 integral dump (for the library's own file reader), byte for byte
    # and the reference's HDF5 integral files (Hamiltonian::read) with the
    # overlap its RHF test loads next to one of them (TestRHF.cpp:29-30) and
    # the two orbital transformations its Hamiltonian test applies to them
    # (TestHam.cpp:84,103)
    for src in ("GKCI_Be_cntSp_cc-pVDZ_0100.txt", "fci_beh2_sto-3gdeterminants.dat", "beh2_sto3g_1_5.dat",
                "atomicintegralsbeh2_sto3g.h5", "atomicintegralsorthonbeh2_sto3g.h5", "overlap.txt",
                "unitary-moorthon.h5", "unitary-mounit_ao_to_mo.txt"):
        shutil.copyfile(os.path.join(REF, src), os.path.join(OUT, src))
        print(src)
    return 0


if __name__ == "__main__":
    sys.exit(main())
