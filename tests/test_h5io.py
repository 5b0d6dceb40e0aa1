"""HDF5 files of the reference (host only, libciflow_h5.so).

Integrals -- Hamiltonian::read / save (Hamiltonian.cpp:322-495) with the
symmetry-packed TwoIndex / FourIndex blocks (TwoIndex.cpp:82-202,
FourIndex.cpp:57-151 and :317-433) -- are pinned by the reference's own files
and answers:
* atomicintegralsbeh2_sto3g.h5 (D2h, non-orthogonal atomic orbitals) read
  with overlap.txt gives the RHF energy of TestRHF.cpp:47,
  -15.12681016931100863 Eh;
* the same file equals the C1 text dump beh2atomic.dat (read by the dump
  parser of make_fixtures.py) after the orbital change to symmetry-adapted
  hydrogen combinations, element for element;
* atomicintegralsorthonbeh2_sto3g.h5 (orthonormalised atomic orbitals) gives
  the FCI energy of the MO-basis dump beh2_sto3g_1_5.dat (FCI is invariant
  under orthonormal orbital changes);
* the packed element count equals the number of orbits of the eight-fold
  permutation group over symmetry-allowed index quadruples (brute force),
  and files written here hold the same packed arrays, byte for byte, as the
  reference-written fixture.
Orbital transformations -- UnitaryMatrix::loadU / saveU
(UnitaryMatrix.cpp:469-526) -- reproduce TestHam.cpp:35-115: either file
turned into MO integrals by the reference's transformation gives the DOCI
energy -15.225362047972723 Eh.  CRS files -- SparseMatrix_CRS::write_to_file /
read_from_file (SparseMatrix_CRS.cpp:484-661) -- round-trip the oracle's
upper CRS.
"""
import itertools
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from ciflow_amd import h5

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
AO_H5 = os.path.join(GOLD, "atomicintegralsbeh2_sto3g.h5")
ORTHON_H5 = os.path.join(GOLD, "atomicintegralsorthonbeh2_sto3g.h5")
E_RHF = -15.12681016931100863  # TestRHF.cpp:47
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

pytestmark = pytest.mark.skipif(not h5.available(), reason="libciflow_h5.so not built (no HDF5 C library)")

PERMS = ((0, 1, 2, 3), (1, 0, 3, 2), (2, 1, 0, 3), (1, 2, 3, 0), (0, 3, 2, 1), (3, 0, 1, 2), (2, 3, 0, 1),
         (3, 2, 1, 0))


def read_overlap(path, irreps):
    """Hamiltonian::load_overlap (Hamiltonian.cpp:524-575): one square block
    per irrep, in irrep order, placed one after the other on the diagonal."""
    L = len(irreps)
    S = np.zeros((L, L))
    lines = open(path).read().splitlines()
    tot, k = 0, 0
    while k < len(lines):
        if lines[k].startswith("irrep_"):
            n = int((np.asarray(irreps) == int(lines[k].split("_")[1])).sum())
            for r in range(n):
                S[tot + r, tot:tot + n] = [float(x) for x in lines[k + 1 + r].split()]
            tot, k = tot + n, k + n + 1
        else:
            k += 1
    assert tot == L
    return S


def rhf_energy(h, V, S, nocc, econst):
    """Closed-shell Roothaan-Hall SCF; (ij|kl) = V[i,k,j,l] in physics notation."""
    import scipy.linalg as sl

    J = lambda P: np.einsum("kl,ikjl->ij", P, V)  # noqa: E731
    K = lambda P: np.einsum("kl,ijkl->ij", P, V)  # noqa: E731
    P = np.zeros_like(h)
    E_old = None
    for _ in range(500):
        e, C = sl.eigh(h + J(P) - 0.5 * K(P), S)
        P = 2.0 * C[:, :nocc] @ C[:, :nocc].T
        E = 0.5 * np.sum(P * (2.0 * h + J(P) - 0.5 * K(P))) + econst
        if E_old is not None and abs(E - E_old) < 1e-13:
            return E
        E_old = E
    raise AssertionError("SCF did not converge")


def orbit_count(irreps):
    """Number of symmetry-distinct elements: orbits of the eight permutations
    of <ij|kl> over the quadruples with irr_i ^ irr_j == irr_k ^ irr_l."""
    L = len(irreps)
    seen = set()
    for q in itertools.product(range(L), repeat=4):
        if irreps[q[0]] ^ irreps[q[1]] != irreps[q[2]] ^ irreps[q[3]]:
            continue
        seen.add(min(tuple(q[p] for p in perm) for perm in PERMS))
    return len(seen)


def random_integrals(irreps, seed):
    """Random h and eight-fold symmetric physics V that respect the point group."""
    L = len(irreps)
    rng = np.random.default_rng(seed)
    h = rng.standard_normal((L, L))
    h = h + h.T
    ir = np.asarray(irreps)
    h[ir[:, None] != ir[None, :]] = 0.0
    V = np.zeros((L,) * 4)
    for q in itertools.product(range(L), repeat=4):
        if ir[q[0]] ^ ir[q[1]] != ir[q[2]] ^ ir[q[3]] or V[q] != 0.0:
            continue
        v = rng.standard_normal()
        for perm in PERMS:
            V[tuple(q[p] for p in perm)] = v
    return h, V


def h5dump_binary(path, dataset, tmp_path):
    out = os.path.join(tmp_path, "dump.bin")
    subprocess.run([H5DUMP, "-d", dataset, "-b", "LE", "-o", out, path], check=True, capture_output=True)
    return open(out, "rb").read()


def h5dump_text(path, obj, tmp_path, attribute=False):
    r = subprocess.run([H5DUMP, "-a" if attribute else "-d", obj, path], check=True, capture_output=True, text=True)
    return r.stdout


def test_library_exports_every_header_symbol():
    import re

    txt = open(os.path.join(os.path.dirname(HERE), "include", "ciflow_amd_h5.h")).read()
    syms = sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(cf_h5_\w+)\s*\(", txt, re.M)))
    assert len(syms) == 7
    missing = [s for s in syms if not hasattr(h5.lib(), s)]
    assert not missing, missing


# ---------------------------------------------------------------------------
# integrals

def test_reads_fixture_header():
    d = h5.read_hamiltonian(AO_H5)
    t = np.load(os.path.join(GOLD, "beh2_atomic.npz"))
    assert (d["L"], d["nalpha"], d["nbeta"], d["ngroup"]) == (7, 3, 3, 7)
    assert list(d["irreps"]) == [0, 0, 0, 5, 5, 6, 7]
    assert d["overlap"] is None
    assert abs(d["econst"] - float(t["econst"])) <= 1e-15


def test_fixture_rhf_energy_matches_reference_test():
    d = h5.read_hamiltonian(AO_H5)
    S = read_overlap(os.path.join(GOLD, "overlap.txt"), d["irreps"])
    E = rhf_energy(d["h"], d["V"], S, 3, d["econst"])
    assert abs(E - E_RHF) <= 1e-9, E


def test_fixture_is_the_text_dump_in_symmetry_adapted_orbitals():
    """Orbitals of the file: Be 1s, Be 2s, (H1+H2)/sqrt2 (Ag); Be 2p_z,
    (H1-H2)/sqrt2 (B1u); Be 2p_x, Be 2p_y (B2u, B3u), as combinations of the
    text dump's Be 1s, 2s, 2p_z, 2p_x, 2p_y, H1 1s, H2 1s."""
    d = h5.read_hamiltonian(AO_H5)
    t = np.load(os.path.join(GOLD, "beh2_atomic.npz"))
    r = np.sqrt(0.5)
    U = np.zeros((7, 7))
    U[0, 0] = U[1, 1] = U[3, 2] = U[5, 3] = U[6, 4] = 1.0
    U[2, 5] = U[2, 6] = U[4, 5] = r
    U[4, 6] = -r
    h = U @ t["h"] @ U.T
    V = np.einsum("ai,bj,ck,dl,ijkl->abcd", U, U, U, U, t["V"], optimize=True)
    assert np.abs(h - d["h"]).max() <= 1e-12
    assert np.abs(V - d["V"]).max() <= 1e-12
    # the comparison can fail: the same change with one orbital's sign flipped does
    U2 = U.copy()
    U2[4] *= -1.0
    assert np.abs(np.einsum("ai,bj,ck,dl,ijkl->abcd", U2, U2, U2, U2, t["V"], optimize=True) - d["V"]).max() > 1e-3


def test_orthonormal_fixture_gives_mo_basis_fci_energy():
    from oracle import oracle as O

    d = h5.read_hamiltonian(ORTHON_H5)
    assert (d["L"], d["nalpha"], d["nbeta"]) == (7, 3, 3)
    gold = json.load(open(os.path.join(GOLD, "golden.json")))["beh2"]["fci_energy"]
    ref = O.build_crs("FCI", d["h"], d["V"], 3, 3)
    E, _, _ = O.lanczos(O.full_from_upper(ref))
    assert abs(E + d["econst"] - gold) <= 1e-9


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_packed_count_is_the_orbit_count(tmp_path):
    d = h5.read_hamiltonian(AO_H5)
    assert orbit_count(list(d["irreps"])) == 106
    assert "( 106 )" in h5dump_text(AO_H5, "/FourIndex/FourIndexObject/Matrix elements", tmp_path)
    for ngroup, irreps in ((7, [0, 0, 0, 5, 5, 6, 7]), (7, [3, 0, 7, 1, 0, 3, 6, 2]), (5, [0, 1, 2, 3, 0, 2]),
                           (0, [0] * 5), (4, [1, 1, 3, 2, 3])):
        h, V = random_integrals(irreps, 11)
        path = os.path.join(tmp_path, f"g{ngroup}_{len(irreps)}.h5")
        h5.save_hamiltonian(path, h, V, 2, 2, 0.5, ngroup, irreps)
        n = orbit_count(irreps)
        assert f"( {n} )" in h5dump_text(path, "/FourIndex/FourIndexObject/Matrix elements", tmp_path), irreps
        assert f"(0): {n}" in h5dump_text(path, "/FourIndex/MetaData/IrrepSizes/theTotalSize", tmp_path, True)


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_written_file_matches_reference_written_bytes(tmp_path):
    for src in (AO_H5, ORTHON_H5):
        d = h5.read_hamiltonian(src)
        out = os.path.join(tmp_path, "copy.h5")
        h5.save_hamiltonian(out, d["h"], d["V"], d["nalpha"], d["nbeta"], d["econst"], d["ngroup"], d["irreps"])
        for ds in ["/FourIndex/FourIndexObject/Matrix elements", "/TwoIndex/MetaData/IrrepSizes",
                   "/FourIndex/MetaData/IrrepSizes", "/Data/orb2irrep", "/Data/Econst", "/Data/L", "/Data/nGroup"] + \
                [f"/TwoIndex/TwoIndex{I}/Matrix elements" for I in (0, 5, 6, 7)]:
            assert h5dump_binary(out, ds, tmp_path) == h5dump_binary(src, ds, tmp_path), (src, ds)


@pytest.mark.parametrize("ngroup,irreps", [(0, [0] * 6), (5, [2, 0, 1, 3, 0, 0, 2]), (7, [7, 0, 4, 4, 1, 0, 6, 3]),
                                           (1, [1, 1, 0])])
def test_round_trip_is_exact(tmp_path, ngroup, irreps):
    h, V = random_integrals(irreps, 5 + len(irreps))
    L = len(irreps)
    S = np.eye(L) + 0.01 * np.arange(L * L).reshape(L, L)
    path = os.path.join(tmp_path, "rt.h5")
    h5.save_hamiltonian(path, h, V, 3, 2, -1.25, ngroup, irreps, overlap=S)
    d = h5.read_hamiltonian(path)
    assert (d["L"], d["nalpha"], d["nbeta"], d["econst"], d["ngroup"]) == (L, 3, 2, -1.25, ngroup)
    assert list(d["irreps"]) == irreps
    assert np.array_equal(d["h"], h) and np.array_equal(d["V"], V) and np.array_equal(d["overlap"], S)
    # without an overlap the file holds an empty one and reads as none
    h5.save_hamiltonian(path, h, V, 3, 2, -1.25, ngroup, irreps)
    assert h5.read_hamiltonian(path)["overlap"] is None


def test_save_refuses_what_the_packed_format_cannot_hold(tmp_path):
    irreps = [0, 1, 0, 1]
    h, V = random_integrals(irreps, 3)
    path = os.path.join(tmp_path, "bad.h5")
    bad = V.copy()
    for perm in PERMS:  # irr 0^0 != 0^1, with all eight permutations
        bad[tuple((0, 0, 0, 1)[p] for p in perm)] = 0.3
    with pytest.raises(h5.H5Error, match="point group"):
        h5.save_hamiltonian(path, h, bad, 1, 1, 0.0, 2, irreps)
    bad_h = h.copy()
    bad_h[0, 1] = bad_h[1, 0] = 0.1
    with pytest.raises(h5.H5Error, match="irreps"):
        h5.save_hamiltonian(path, bad_h, V, 1, 1, 0.0, 2, irreps)
    with pytest.raises(h5.H5Error, match="out of range"):
        h5.save_hamiltonian(path, h, V, 1, 1, 0.0, 2, [0, 2, 0, 1])  # C2 has 2 irreps
    asym = V.copy()
    asym[0, 1, 0, 1] += 0.5  # breaks <ij|kl> = <kj|il> and the other permutations
    with pytest.raises(h5.H5Error, match="eight-fold"):
        h5.save_hamiltonian(path, h, asym, 1, 1, 0.0, 2, irreps)
    hs = h.copy()
    hs[0, 2] += 0.1
    with pytest.raises(h5.H5Error, match="not symmetric"):
        h5.save_hamiltonian(path, hs, V, 1, 1, 0.0, 2, irreps)
    with pytest.raises(h5.H5Error):
        h5.read_hamiltonian(os.path.join(tmp_path, "missing.h5"))


# ---------------------------------------------------------------------------
# orbital transformations (TestHam.cpp:35-115)

E_DOCI = -15.225362047972723  # TestHam.cpp:42,93,110 (TestDOCI.cpp:31)


def read_unitary_text(path, irreps):
    """UnitaryMatrix::load_unitary text form (UnitaryMatrix.cpp:572-610): after
    "CIFlowTransformation", per irrep with orbitals a header line holding
    "irrep" and the block's rows (new orbitals)."""
    L = len(irreps)
    U = np.zeros((L, L))
    g = [[p for p in range(L) if irreps[p] == I] for I in range(8)]
    lines = open(path).read().splitlines()
    k = next(i for i, ln in enumerate(lines) if "CIFlowTransformation" in ln)
    for I in range(8):
        if not g[I]:
            continue
        while "irrep" not in lines[k].lower():
            k += 1
        for a in range(len(g[I])):
            U[g[I][a], g[I]] = [float(x) for x in lines[k + 1 + a].split()]
        k += len(g[I]) + 1
    return U


def doci_energy(h, V, econst):
    from oracle import oracle as O

    E, _, _ = O.lanczos(O.full_from_upper(O.build_crs("DOCI", h, V, 3, 3)))
    return E + econst


def rotate(h, V, U):
    return U @ h @ U.T, np.einsum("ai,bj,ck,dl,ijkl->abcd", U, U, U, U, V, optimize=True)


@pytest.mark.parametrize("ham,unit", [("atomicintegralsorthonbeh2_sto3g.h5", "unitary-moorthon.h5"),
                                      ("atomicintegralsbeh2_sto3g.h5", "unitary-mounit_ao_to_mo.txt")])
def test_ao_to_mo_transformation_gives_doci_kat(ham, unit):
    d = h5.read_hamiltonian(os.path.join(GOLD, ham))
    upath = os.path.join(GOLD, unit)
    U = h5.read_unitary(upath, d["irreps"], d["ngroup"]) if unit.endswith(".h5") else \
        read_unitary_text(upath, d["irreps"])
    h, V = rotate(d["h"], d["V"], U)
    assert abs(doci_energy(h, V, d["econst"]) - E_DOCI) <= 1e-9
    # the transpose (columns as new orbitals) is a different Hamiltonian
    h2, V2 = rotate(d["h"], d["V"], U.T)
    assert abs(doci_energy(h2, V2, d["econst"]) - E_DOCI) > 1e-6


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_unitary_round_trip_and_reference_bytes(tmp_path):
    d = h5.read_hamiltonian(ORTHON_H5)
    src = os.path.join(GOLD, "unitary-moorthon.h5")
    U = h5.read_unitary(src, d["irreps"], d["ngroup"])
    assert np.allclose(U @ U.T, np.eye(7), atol=1e-8)  # orthonormal to orthonormal orbitals
    out = os.path.join(tmp_path, "u.h5")
    h5.save_unitary(out, U, d["irreps"], d["ngroup"])
    assert np.array_equal(h5.read_unitary(out, d["irreps"], d["ngroup"]), U)
    for I in (0, 5, 6, 7):
        assert h5dump_binary(out, f"/Data/irrep_{I}", tmp_path) == h5dump_binary(src, f"/Data/irrep_{I}", tmp_path)
    bad = U.copy()
    bad[0, 3] = 0.5
    with pytest.raises(h5.H5Error, match="irreps"):
        h5.save_unitary(out, bad, d["irreps"], d["ngroup"])


# ---------------------------------------------------------------------------
# CRS

@pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")
def test_crs_round_trip_and_layout(tmp_path):
    from oracle import oracle as O

    b = np.load(os.path.join(GOLD, "beh2_sto3g.npz"))
    fci = O.build_crs("FCI", b["h"], b["V"], 3, 3)
    doci = O.build_crs("DOCI", b["h"], b["V"], 3, 3)
    path = os.path.join(tmp_path, "crs.h5")
    h5.write_crs(path, "fci", fci.row, fci.col, fci.data)
    h5.write_crs(path, "doci", doci.row, doci.col, doci.data, append=True)
    for name, m in (("fci", fci), ("doci", doci)):
        row, col, data = h5.read_crs(path, name)
        assert np.array_equal(row, np.asarray(m.row, dtype=np.uint32))
        assert np.array_equal(col, np.asarray(m.col, dtype=np.uint32))
        assert np.array_equal(data, np.asarray(m.data))
    txt = h5dump_text(path, "/fci/row", tmp_path)
    assert "H5T_STD_U64LE" in txt and f"( {len(fci.row)} )" in txt
    assert f"(0): {len(fci.data)}" in h5dump_text(path, "/fci/data/size", tmp_path, True)
    assert f"(0): {len(fci.data)}" in h5dump_text(path, "/fci/col/size", tmp_path, True)
    assert "H5T_STD_U64LE" in h5dump_text(path, "/fci/n", tmp_path)
    # an empty matrix and the error paths
    h5.write_crs(path, "empty", np.zeros(4, dtype=np.uint32), np.zeros(0, dtype=np.uint32), np.zeros(0), append=True)
    row, col, data = h5.read_crs(path, "empty")
    assert list(row) == [0, 0, 0, 0] and len(col) == 0 and len(data) == 0
    with pytest.raises(h5.H5Error):
        h5.read_crs(path, "absent")
    with pytest.raises(h5.H5Error, match="row"):
        h5.write_crs(path, "broken", np.array([0, 5], dtype=np.uint32), np.zeros(2, dtype=np.uint32), np.zeros(2),
                     append=True)
