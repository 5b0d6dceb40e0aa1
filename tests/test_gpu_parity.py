"""GPU parity tests: the HIP library through its C ABI against the oracle.

Bar (from the reference's contract): determinant order, CRS pattern and signs
bit-exact; values bit-exact here as well (same formulas, same summation
order); energies within 1e-10 Eh; RDM elements within 1e-9; FP32-value solve
with FP64 polish within 1e-8 Eh.
"""
import json
import os
import sys

import numpy as np
import pytest

import ciflow_amd as CA
from oracle import bruteforce as BF
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def golden():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ctx():
    return CA.default_context()


def fixture(name):
    d = np.load(os.path.join(GOLD, name))
    return d["h"], d["V"], int(d["nalpha"]), int(d["nbeta"]), float(d["econst"])


def redbcs(L, eps, g):
    h = np.diag(np.asarray(eps, dtype=float))
    V = np.zeros((L, L, L, L))
    for i in range(L):
        for j in range(L):
            V[i, i, j, j] = g
    return h, V


# ---- (a) ranking / unranking --------------------------------------------------
@pytest.mark.parametrize("L,n", [(8, 4), (12, 6), (16, 8), (1, 1), (5, 0), (20, 10)])
def test_unrank_matches_gosper_order(ctx, L, n):
    say("unrank", L, n)
    ref = O.enumerate_strings(L, n)
    got = ctx.unrank(L, n, 0, len(ref))
    assert np.array_equal(got, ref)
    r = ctx.rank(L, n, ref)
    assert np.array_equal(r, np.arange(len(ref), dtype=np.uint32))


def test_rank_unrank_large_space(ctx):
    L, n = 24, 12
    d = O.binom(L, n)
    ref = O.enumerate_strings(L, n)
    assert np.array_equal(ctx.unrank(L, n, 0, d), ref)
    idx = np.unique(np.random.default_rng(0).integers(0, d, 20000))
    assert np.array_equal(ctx.rank(L, n, ref[idx]), idx.astype(np.uint32))
    assert np.array_equal(ctx.unrank(L, n, d - 1000, 1000), ref[d - 1000:])


# ---- (b) CRS build -----------------------------------------------------------
CASES = {
    "fci_L4_2a2b": ("FCI", lambda: (*CA.gen_integrals(4, 5), 2, 2)),
    "fci_L6_3a3b": ("FCI", lambda: (*CA.gen_integrals(6, 7), 3, 3)),
    "fci_L7_2a2b": ("FCI", lambda: (*CA.gen_integrals(7, 2), 2, 2)),
    "fci_L8_4a4b_config1": ("FCI", lambda: (*CA.gen_integrals(8, 0), 4, 4)),
    "fci_beh2": ("FCI", lambda: fixture("beh2_sto3g.npz")[:4]),
    "doci_beh2": ("DOCI", lambda: fixture("beh2_sto3g.npz")[:4]),
    "doci_L10_5": ("DOCI", lambda: (*CA.gen_integrals(10, 11), 5, 5)),
    "doci_redbcs_L8": ("DOCI", lambda: (*redbcs(8, np.arange(8.0), -0.5), 4, 4)),
}


def build_pair(kind, h, V, na, nb, prec=CA.CF_F64):
    ham = CA.Hamiltonian(h, V, na, nb)
    cim = (CA.FCI if kind == "FCI" else CA.DOCI)(ham, value_prec=prec)
    ref = O.build_crs(kind, h, V, na, nb)
    return ham, cim, ref


@pytest.mark.parametrize("case", list(CASES))
def test_crs_upper_bit_exact(case):
    kind, make = CASES[case]
    h, V, na, nb = make()
    say("crs", case)
    ham, cim, ref = build_pair(kind, h, V, na, nb)
    row, col, val = cim.get_mat().export_upper()
    assert np.array_equal(row, ref.row)
    assert np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.data)  # bit-exact values, hence signs
    full = O.full_from_upper(ref)
    frow, fcol, fval = cim.get_mat().export_full()
    assert np.array_equal(frow, full.row) and np.array_equal(fcol, full.col) and np.array_equal(fval, full.data)


def test_config1_digest_matches_golden():
    g = golden()["config1"]
    h, V = CA.gen_integrals(8, 0)
    cim = CA.FCI(CA.Hamiltonian(h, V, 4, 4))
    n, nnz_full, nnz_up = cim.get_mat().info()
    assert (n, nnz_up, nnz_full) == (4900, 886900, 1768900)


def test_upper_export_with_32_bit_row_pointer():
    """cf_crs_export_upper_u32: the reference's unsigned-int row pointer
    (SparseMatrix_CRS.h:111-112), equal to the 64-bit export, whole and in a
    row range."""
    h, V = CA.gen_integrals(8, 0)
    m = CA.FCI(CA.Hamiltonian(h, V, 4, 4)).get_mat()
    for r0, r1 in [(0, m.getn()), (1234, 3001)]:
        row, col, val = m.export_upper(r0, r1)
        row32, col32, val32 = m.export_upper_u32(r0, r1)
        assert row32.dtype == np.uint32 and np.array_equal(row32.astype(np.uint64), row)
        assert np.array_equal(col32, col) and np.array_equal(val32, val)


def test_kat_energies_from_the_integral_file():
    """Hamiltonian(filename): the reference's BeH2 dump read by the library,
    DOCI KAT -15.225362047972723 (TestDOCI.cpp:31) and the FCI energy."""
    ham = CA.Hamiltonian.from_file(os.path.join(GOLD, "beh2_sto3g_1_5.dat"))
    d = CA.DOCI(ham)
    d.solve()
    assert abs(d.get_ci_energy() - (-15.225362047972723)) < 1e-10
    f = CA.FCI(ham)
    f.solve()
    assert abs(f.get_ci_energy() - golden()["beh2"]["fci_energy"]) < 1e-10


@pytest.mark.skipif(not CA.h5.available(), reason="libciflow_h5.so not built (no HDF5 C library)")
def test_hdf5_integrals_and_crs_file(tmp_path):
    """Hamiltonian::read of the reference's orthonormal-AO BeH2 file (D2h):
    its FCI energy is the MO-basis one (orbital-rotation invariance); the
    GPU CRS written by SparseMatrix_CRS::write_to_file reads back equal to the
    oracle's upper CRS for the same integrals; Hamiltonian::save round-trips;
    the file's transformation to MOs gives the DOCI KAT."""
    path = os.path.join(GOLD, "atomicintegralsorthonbeh2_sto3g.h5")
    ham = CA.Hamiltonian.read(path)
    assert list(ham.irreps) == [0, 0, 0, 5, 5, 6, 7] and ham.ngroup == 7
    f = CA.FCI(ham)
    f.solve()
    assert abs(f.get_ci_energy() - golden()["beh2"]["fci_energy"]) < 1e-10
    out = os.path.join(tmp_path, "crs.h5")
    f.get_mat().write_to_file(out, "fci")
    row, col, val = CA.h5.read_crs(out, "fci")
    ref = O.build_crs("FCI", ham.h, ham.V, 3, 3)
    assert np.array_equal(row, np.asarray(ref.row, dtype=np.uint32))
    assert np.array_equal(col, ref.col) and np.array_equal(val, ref.data)
    saved = os.path.join(tmp_path, "ham.h5")
    ham.save(saved)
    d = CA.h5.read_hamiltonian(saved)
    assert np.array_equal(d["h"], ham.h) and np.array_equal(d["V"], ham.V) and d["econst"] == ham.econst
    # TestHam.cpp:80-93: the reference's transformation to MOs (UnitaryMatrix::loadU),
    # applied on the GPU, gives the DOCI KAT
    U = CA.h5.read_unitary(os.path.join(GOLD, "unitary-moorthon.h5"), ham.irreps, ham.ngroup)
    doci = CA.DOCI(ham.rotate(U))
    doci.solve()
    assert abs(doci.get_ci_energy() - (-15.225362047972723)) < 1e-10


# ---- selected CI over determinant lists (FCI_File) ----------------------------
DET_FILES = {"be_gkci100": ("be_ccpvdz.npz", "GKCI_Be_cntSp_cc-pVDZ_0100.txt"),
             "beh2_full_list": ("beh2_sto3g.npz", "fci_beh2_sto-3gdeterminants.dat")}


def check_det_crs(cim, ref):
    row, col, val = cim.get_mat().export_upper()
    assert np.array_equal(row, ref.row)
    assert np.array_equal(col, ref.col)
    assert np.array_equal(val, ref.data)  # bit-exact values, hence signs
    full = O.full_from_upper(ref)
    frow, fcol, fval = cim.get_mat().export_full()
    assert np.array_equal(frow, full.row) and np.array_equal(fcol, full.col) and np.array_equal(fval, full.data)


@pytest.mark.parametrize("case", list(DET_FILES))
def test_det_file_crs_bit_exact_and_kat_energy(case):
    """FCI_File from the reference's determinant files: the upper CRS equals the
    oracle's FCI_File::construct_CI_matrix bit for bit, and the energy meets the
    KAT (TestDetFile.cpp:32) or, for the full BeH2 list, the FCI energy."""
    say("det file", case)
    npz, dets = DET_FILES[case]
    h, V, na, nb, ec = fixture(npz)
    path = os.path.join(GOLD, dets)
    cim = CA.FCI_File(CA.Hamiltonian(h, V, na, nb, ec), path)
    up, down = O.read_dets(path)
    assert np.array_equal(cim.up, up) and np.array_equal(cim.down, down)
    check_det_crs(cim, O.build_crs_dets(h, V, na, nb, up, down))
    info = cim.solve()
    assert info.status == 0
    expected = -14.5951673370889 if case == "be_gkci100" else golden()["beh2"]["fci_energy"]
    assert abs(cim.get_ci_energy() - expected) < 1e-10


@pytest.mark.parametrize("L,na,nb,count,seed", [(8, 4, 4, 3000, 1), (9, 5, 3, 2000, 2), (7, 2, 2, 441, 3)])
def test_det_list_random_subset_bit_exact(L, na, nb, count, seed):
    """Shuffled subsets of an FCI space (the whole space for L=7 2a2b): CRS
    bit-exact, FP64 energy against dense eigh of the oracle matrix, FP32-value
    solve with FP64 polish within 1e-8."""
    say("det subset", L, na, nb, count)
    h, V = CA.gen_integrals(L, seed)
    A = O.enumerate_strings(L, na)
    B = O.enumerate_strings(L, nb)
    rng = np.random.default_rng(seed)
    pick = rng.permutation(len(A) * len(B))[:count]
    up = A[pick // len(B)]
    down = B[pick % len(B)]
    cim = CA.FCI_File(CA.Hamiltonian(h, V, na, nb), (up, down), value_prec=CA.CF_F32)
    ref = O.build_crs_dets(h, V, na, nb, up, down)
    check_det_crs(cim, ref)
    w = np.linalg.eigvalsh(O.dense_from_upper(ref))[0]
    cim.solve()
    assert abs(cim.get_ci_energy() - w) < 1e-10
    cim.solve(precision=CA.CF_F32)
    assert abs(cim.get_ci_energy() - w) < 1e-8


def test_det_list_block_with_repeats_bit_exact():
    """A CAS-like block (the first 40 alpha strings of FCI L=10 5a5b with every
    beta string, 10,080 determinants, shuffled) with three determinants listed
    twice: the CRS built by excitation lookup in the sorted keys equals the
    oracle's all-pairs FCI_File::construct_CI_matrix bit for bit."""
    L, na, nb = 10, 5, 5
    h, V = CA.gen_integrals(L, 21)
    A = O.enumerate_strings(L, na)
    B = O.enumerate_strings(L, nb)
    ia = np.repeat(np.arange(40), len(B))
    ib = np.tile(np.arange(len(B)), 40)
    perm = np.random.default_rng(21).permutation(len(ia))
    up, down = A[ia[perm]], B[ib[perm]]
    up = np.concatenate([up, up[[5, 777, 9000]]])
    down = np.concatenate([down, down[[5, 777, 9000]]])
    say("det block", len(up))
    cim = CA.FCI_File(CA.Hamiltonian(h, V, na, nb), (up, down))
    check_det_crs(cim, O.build_crs_dets(h, V, na, nb, up, down))


def test_det_list_beyond_32_orbitals_bit_exact():
    """L = 34 (keys no longer fit 32 bits per spin): the all-pairs build, bit-exact."""
    L, na, nb = 34, 2, 2
    h, V = CA.gen_integrals(L, 4)
    rng = np.random.default_rng(4)
    A = O.enumerate_strings(L, na)
    pick = rng.choice(len(A) * len(A), 300, replace=False)
    up, down = A[pick // len(A)], A[pick % len(A)]
    cim = CA.FCI_File(CA.Hamiltonian(h, V, na, nb), (up, down))
    check_det_crs(cim, O.build_crs_dets(h, V, na, nb, up, down))



def _cas_block(L, n_alpha, seed, repeats=()):
    """The first n_alpha alpha strings times every beta string (half filling),
    shuffled, with the listed determinants appended again."""
    na = L // 2
    S = O.enumerate_strings(L, na)
    ia = np.repeat(np.arange(n_alpha), len(S))
    ib = np.tile(np.arange(len(S)), n_alpha)
    perm = np.random.default_rng(seed).permutation(len(ia))
    up, down = S[ia[perm]], S[ib[perm]]
    rep = list(repeats)
    return np.concatenate([up, up[rep]]), np.concatenate([down, down[rep]])


@pytest.mark.parametrize("case", ["beh2_full_list", "be_gkci100", "cas_block_L10", "fci_L8_4a4b_config1",
                                  "doci_L10_5", "cas_block_L12"])
def test_sliced_spmv_matches_oracle(case):
    """Kernel 7 (rows in chunks of 64 with interleaved entries, each chunk
    padded to its longest row; launches with few chunks split each into quad
    ranges added by a second pass): every waves x unroll shape against the
    oracle's product (FP64 values, 1e-12 relative; for the 31,416-determinant
    block, whose all-pairs oracle build is slow, the vectorised row gather on
    the same values, 1e-13) and against the scalar row gather on the same FP32
    values; ragged rows (a shuffled CAS-like block with repeated determinants,
    D2h zeros dropped), last chunks with fewer than 64 rows (100 and 1,225
    determinants), repeated products bitwise equal."""
    say("sliced spmv", case)
    if case in DET_FILES:
        npz, dets = DET_FILES[case]
        h, V, na, nb, ec = fixture(npz)
        up, down = O.read_dets(os.path.join(GOLD, dets))
        cim = CA.FCI_File(CA.Hamiltonian(h, V, na, nb, ec), (up, down), value_prec=CA.CF_F32)
        ref = O.build_crs_dets(h, V, na, nb, up, down)
    elif case == "cas_block_L10":
        h, V = CA.gen_integrals(10, 23)
        up, down = _cas_block(10, 37, 23, repeats=(3, 4000, 9000))
        cim = CA.FCI_File(CA.Hamiltonian(h, V, 5, 5), (up, down), value_prec=CA.CF_F32)
        ref = O.build_crs_dets(h, V, 5, 5, up, down)
    elif case == "cas_block_L12":
        h, V = CA.gen_integrals(12, 24)
        up, down = _cas_block(12, 34, 24)
        cim = CA.FCI_File(CA.Hamiltonian(h, V, 6, 6), (up, down), value_prec=CA.CF_F32)
        ref = None
    else:
        kind, make = CASES[case]
        h, V, na, nb = make()
        ham, cim, ref = build_pair(kind, h, V, na, nb, prec=CA.CF_F32)
    m = cim.get_mat()
    n = cim.get_dim()
    x = np.random.default_rng(11).standard_normal(n)
    m.set_spmv_kernel(3)
    y_ref = O.mvprod_sym(ref, x) if ref is not None else m.mvprod(x)
    tol = 1e-12 if ref is not None else 1e-13
    m.set_spmv_kernel(1)
    y1_32 = m.mvprod(x, CA.CF_F32)
    m.set_spmv_kernel(7)  # builds the sliced copy
    assert m.spmv_bytes(CA.CF_F64)[0] == 7
    shapes = [(w, u) for w in (1, 2, 4, 8, 16) for u in (1, 2)] + [(4, 4), (8, 4), (16, 4)]
    for w, u in shapes:
        m.set_spmv_shape(CA.CF_F64, 7, w, u)
        y = m.mvprod(x, CA.CF_F64)
        assert np.abs(y - y_ref).max() <= tol * np.abs(y_ref).max(), (case, w, u)
        assert np.array_equal(m.mvprod(x, CA.CF_F64), y)
        m.set_spmv_shape(CA.CF_F32, 7, w, u)
        y32 = m.mvprod(x, CA.CF_F32)
        assert np.abs(y32 - y1_32).max() <= 1e-13 * np.abs(y1_32).max(), (case, w, u)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 7, 3, 1)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 7, 4, 3)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 7, 2, 4)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_kernel(7, 32)


def test_sliced_spmv_ragged_random_list():
    """Kernel 7 on 100,000 random determinants of FCI L=14 7a7b in random order
    (rows of 1 to ~100 entries, so the chunks' padding is large and the
    launches split their chunks): every shape against the vectorised CRS row
    gather on the same values (1e-13 relative), FP64 and FP32.  The list is
    not in (up, down) order, so the automatic choice stays the row gather."""
    L, na = 14, 7
    h, V = CA.gen_integrals(L, 31)
    S = O.enumerate_strings(L, na)
    pick = np.random.default_rng(31).choice(len(S) * len(S), 100000, replace=False)
    up, down = S[pick // len(S)], S[pick % len(S)]
    m = CA.FCI_File(CA.Hamiltonian(h, V, na, na), (up, down), value_prec=CA.CF_F32).get_mat()
    assert m.spmv_bytes(CA.CF_F64)[0] != 7 and m.spmv_bytes(CA.CF_F32)[0] != 7
    order = np.lexsort((down, up))
    sorted_m = CA.FCI_File(CA.Hamiltonian(h, V, na, na), (up[order], down[order])).get_mat()
    assert sorted_m.spmv_bytes(CA.CF_F64)[0] == 7
    del sorted_m
    x = np.random.default_rng(32).standard_normal(len(up))
    for prec in (CA.CF_F64, CA.CF_F32):
        m.set_spmv_shape(prec, 3, 16, 2)
        y3 = m.mvprod(x, prec)
        for w, u in ((16, 2), (16, 4), (8, 1), (4, 2), (1, 1)):
            m.set_spmv_shape(prec, 7, w, u)
            y7 = m.mvprod(x, prec)
            assert np.abs(y7 - y3).max() <= 1e-13 * np.abs(y3).max(), (prec, w, u)
            assert np.array_equal(m.mvprod(x, prec), y7)


def test_sliced_spmv_follows_set_ham_and_autotune():
    """After new integrals on a resident determinant-list or small-alpha FCI
    matrix the sliced copy is rebuilt from the new CRS (a fresh build's
    product, 1e-13), also when the autotuner is the first to read it; the
    autotuner times the sliced shapes and installs the fastest, whose product
    agrees."""
    L, na = 8, 4
    h, V = CA.gen_integrals(L, 12)
    h2, V2 = CA.gen_integrals(L, 13)
    up, down = _cas_block(L, 30, 12)
    sel = CA.FCI_File(CA.Hamiltonian(h, V, na, na), (up, down))
    m = sel.get_mat()
    x = np.random.default_rng(12).standard_normal(len(up))
    m.set_spmv_kernel(7, 2)
    y_old = m.mvprod(x)
    sel.set_ham(CA.Hamiltonian(h2, V2, na, na))
    ref = CA.FCI_File(CA.Hamiltonian(h2, V2, na, na), (up, down)).get_mat()
    y_new = ref.mvprod(x)
    y = sel.get_mat().mvprod(x)
    assert sel.get_mat().spmv_bytes(CA.CF_F64)[0] == 7
    assert np.abs(y - y_new).max() <= 1e-13 * np.abs(y_new).max()
    assert np.abs(y_old - y_new).max() > 1e-3
    tab = ref.autotune(CA.CF_F64, reps=2)
    assert 7 in {k for k, _, _, _ in tab} and 3 in {k for k, _, _, _ in tab}
    best = min(tab, key=lambda t: t[3])
    assert ref.spmv_bytes(CA.CF_F64)[0] == best[0]
    assert np.abs(ref.mvprod(x) - y_new).max() <= 1e-13 * np.abs(y_new).max()
    # the autotuner on the resident matrix refills its stale sliced copy
    tab = sel.get_mat().autotune(CA.CF_F64, reps=2)
    assert 7 in {k for k, _, _, _ in tab}
    assert np.abs(sel.get_mat().mvprod(x) - y_new).max() <= 1e-13 * np.abs(y_new).max()
    # FCI with too few alpha strings for the position-major copy (126) gets the
    # sliced copy at build; new integrals refill it
    a9, b9 = CA.gen_integrals(9, 5), CA.gen_integrals(9, 6)
    f = CA.FCI(CA.Hamiltonian(*a9, 4, 4))
    assert f.get_mat().spmv_bytes(CA.CF_F64)[0] == 7
    f.set_ham(CA.Hamiltonian(*b9, 4, 4))
    x9 = np.random.default_rng(13).standard_normal(f.get_dim())
    y9 = CA.FCI(CA.Hamiltonian(*b9, 4, 4)).get_mat().mvprod(x9)
    assert np.abs(f.get_mat().mvprod(x9) - y9).max() <= 1e-13 * np.abs(y9).max()
    f.set_ham(CA.Hamiltonian(*a9, 4, 4))
    f.set_ham(CA.Hamiltonian(*b9, 4, 4))
    tab = f.get_mat().autotune(CA.CF_F64, reps=2)
    assert 7 in {k for k, _, _, _ in tab}
    assert np.abs(f.get_mat().mvprod(x9) - y9).max() <= 1e-13 * np.abs(y9).max()


# ---- CI_Big: determinant lists with the product on the fly --------------------
@pytest.mark.parametrize("case,stored", [("beh2_full_list", 0), ("beh2_full_list", 300),
                                         ("beh2_full_list", 1225), ("be_gkci100", 40)])
def test_ci_big_product_matches_stored_matrix(case, stored):
    """CI_Big::mvprod (CIMethod.cpp:943-984) with no rows, some leading rows
    or all rows stored: equals the oracle's FCI_File matrix product to
    1e-12 relative, and repeats bitwise."""
    npz, dets = DET_FILES[case]
    h, V, na, nb, ec = fixture(npz)
    up, down = O.read_dets(os.path.join(GOLD, dets))
    say("ci_big product", case, stored)
    big = CA.CI_Big(CA.Hamiltonian(h, V, na, nb, ec), (up, down), stored_rows=min(stored, len(up)))
    ref = O.build_crs_dets(h, V, na, nb, up, down)
    for seed in (0, 1):
        x = np.random.default_rng(seed).standard_normal(len(up))
        y = big.mvprod(x)
        y_ref = O.mvprod_sym(ref, x)
        assert np.abs(y - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
        assert np.array_equal(big.mvprod(x), y)


def test_ci_big_kat_energy_and_rdms():
    """CI_Big on the reference's GKCI list (Be cc-pVDZ, TestDetFile.cpp:32):
    the KAT energy with 40 of 100 rows stored and with none; its RDMs equal
    FCI_File's for the same vector."""
    h, V, na, nb, ec = fixture("be_ccpvdz.npz")
    path = os.path.join(GOLD, "GKCI_Be_cntSp_cc-pVDZ_0100.txt")
    ham = CA.Hamiltonian(h, V, na, nb, ec)
    for stored in (40, 0):
        big = CA.CI_Big(ham, path, stored_rows=stored)
        big.solve()
        assert abs(big.get_ci_energy() - (-14.5951673370889)) < 1e-10
    one, two = big.construct_density()
    ref = CA.FCI_File(ham, path)
    ref._eigvec = big.get_eigvec()
    one_r, two_r = ref.construct_density()
    assert np.abs(one - one_r).max() <= 1e-12 and np.abs(two - two_r).max() <= 1e-12


def test_ci_big_larger_list_against_gpu_crs():
    """A 30,000-determinant random subset of FCI L=12 6a6b with duplicates,
    8,000 leading rows stored: the on-the-fly product equals the GPU
    FCI_File CRS product (itself bit-exact against the oracle's pattern) to
    1e-12, and the ground-state energies agree to 1e-10."""
    L, na, nb = 12, 6, 6
    h, V = CA.gen_integrals(L, 9)
    A = O.enumerate_strings(L, na)
    rng = np.random.default_rng(9)
    pick = rng.integers(0, len(A) * len(A), 30000)
    up, down = A[pick // len(A)], A[pick % len(A)]
    ham = CA.Hamiltonian(h, V, na, nb)
    say("ci_big L=12 list", len(up))
    big = CA.CI_Big(ham, (up, down), stored_rows=8000)
    crs = CA.FCI_File(ham, (up, down))
    x = rng.standard_normal(len(up))
    y = big.mvprod(x)
    y_ref = crs.get_mat().mvprod(x)
    assert np.abs(y - y_ref).max() <= 1e-12 * np.abs(y_ref).max()
    big.solve()
    crs.solve()
    assert abs(big.get_ci_energy() - crs.get_ci_energy()) < 1e-10


def test_ci_big_rejects_bad_input():
    h, V = CA.gen_integrals(6, 1)
    ham = CA.Hamiltonian(h, V, 3, 3)
    A = O.enumerate_strings(6, 3)
    with pytest.raises(CA.CIFlowError):
        CA.CI_Big(ham, (A[:5], A[:5]), stored_rows=6)
    f = CA.FCI(ham, build=False)  # not a determinant-list space
    out = CA.vp()
    with pytest.raises(CA.CIFlowError, match="determinant-list"):
        CA._check(CA.lib().cf_dets_sigma_create(f.ptr, 0, CA.C.byref(out)))


def test_det_list_rejects_bad_input():
    h, V = CA.gen_integrals(6, 1)
    ham = CA.Hamiltonian(h, V, 3, 3)
    with pytest.raises(CA.CIFlowError):  # two beta electrons instead of three
        CA.FCI_File(ham, (np.array([7], dtype=np.uint64), np.array([3], dtype=np.uint64)))
    with pytest.raises(CA.CIFlowError):  # orbital 6 outside L = 6
        CA.FCI_File(ham, (np.array([7 << 4], dtype=np.uint64), np.array([7], dtype=np.uint64)))
    cim = CA.FCI_File(ham, (np.array([7, 11, 7], dtype=np.uint64), np.array([7, 7, 7], dtype=np.uint64)))
    assert cim.get_mat().info()[0] == 3
    cim._eigvec = np.array([0.6, 0.8, 0.0])
    with pytest.raises(CA.CIFlowError):  # a repeated determinant has no RDM
        cim.construct_density()


def test_det_file_rdms_match_fci_rdms():
    """TestDensity.cpp: DensFILE on the full BeH2 list equals DensFCI, and the
    density energy equals the CI energy."""
    h, V, na, nb, ec = fixture("beh2_sto3g.npz")
    ham = CA.Hamiltonian(h, V, na, nb, ec)
    f = CA.FCI_File(ham, os.path.join(GOLD, "fci_beh2_sto-3gdeterminants.dat"))
    f.solve()
    one_f, two_f = f.construct_density()
    fci = CA.FCI(ham)
    fci.solve()
    one, two = fci.construct_density()
    _rdm_compare(h.shape[0], one_f, two_f, one, two)
    assert abs(f.get_dens_energy() - f.get_ci_energy()) < 1e-10


@pytest.mark.parametrize("L,na,nb,count,seed", [(6, 3, 2, 150, 4), (7, 2, 2, 200, 5), (5, 2, 2, 100, 6)])
def test_det_list_rdms_match_second_quantisation(L, na, nb, count, seed):
    """Shuffled subsets of a space (the whole space for L=5 2a2b): the dense RDMs
    of the determinant-list path equal <a+a> and <a+a+aa> evaluated with
    explicit operators on the same vector."""
    h, V = CA.gen_integrals(L, seed)
    A = O.enumerate_strings(L, na)
    B = O.enumerate_strings(L, nb)
    pick = np.random.default_rng(seed).permutation(len(A) * len(B))[:count]
    up = A[pick // len(B)]
    down = B[pick % len(B)]
    cim = CA.FCI_File(CA.Hamiltonian(h, V, na, nb), (up, down))
    cim.solve()
    c = cim.get_eigvec()
    g, G = cim.density_dense(c)
    gb, Gb = BF.rdms(c, L, na, nb, dets=[(int(a), int(b)) for a, b in zip(up, down)])
    assert np.max(np.abs(g - gb)) < 1e-12
    assert np.max(np.abs(G - Gb)) < 1e-12



# ---- (c) SpMV ----------------------------------------------------------------
@pytest.mark.parametrize("case", ["fci_L8_4a4b_config1", "doci_L10_5", "fci_beh2"])
def test_mvprod_matches_oracle(case):
    kind, make = CASES[case]
    h, V, na, nb = make()
    ham, cim, ref = build_pair(kind, h, V, na, nb)
    x = np.random.default_rng(1).standard_normal(ref.n)
    y = cim.get_mat().mvprod(x)
    yr = O.mvprod_sym(ref, x)
    assert np.max(np.abs(y - yr)) <= 1e-12 * np.max(np.abs(yr))


# ---- solve --------------------------------------------------------------------
def test_config1_energy():
    g = golden()["config1"]
    h, V = CA.gen_integrals(8, 0)
    cim = CA.FCI(CA.Hamiltonian(h, V, 4, 4))
    info = cim.solve()
    say("config1", info.as_dict())
    assert info.status == 0
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-10
    c = cim.get_eigvec()
    assert abs(np.linalg.norm(c) - 1) < 1e-12


def test_doci_beh2_kat_on_gpu():
    """TestDOCI.cpp:31."""
    h, V, na, nb, ec = fixture("beh2_sto3g.npz")
    cim = CA.DOCI(CA.Hamiltonian(h, V, na, nb, ec))
    cim.solve()
    assert abs(cim.get_ci_energy() - (-15.225362047972723)) <= 1e-12 * 15.23


def test_fci_beh2_energy_and_density_identity():
    """TestDensity.cpp:57: E_dens == E_FCI."""
    h, V, na, nb, ec = fixture("beh2_sto3g.npz")
    cim = CA.FCI(CA.Hamiltonian(h, V, na, nb, ec))
    cim.solve()
    E = cim.get_ci_energy()
    assert abs(E - golden()["beh2"]["fci_energy"]) < 1e-10
    assert abs(cim.get_dens_energy() - E) <= 1e-12 * abs(E) + 1e-12


def test_redbcs_kat_on_gpu():
    """TestRedBCS.cpp:44: E_red = -2.5978142225."""
    h0, V0, na, nb, ec = fixture("psi0_631g.npz")
    L = h0.shape[0]
    h, V = redbcs(L, np.sort(np.diag(h0)), 0.119034351059)
    cim = CA.DOCI(CA.Hamiltonian(h, V, na, nb, 0.0))
    cim.solve()
    assert abs(cim.get_ci_energy() - (-2.5978142225)) <= 1e-9 * 2.6


def test_rci_dropin_with_gpu_mvprod():
    """The reference's arpackDiagonalize loop shape with the GPU mvprod as the
    virtual seam (SparseMatrix_CRS.cpp:441-449)."""
    g = golden()["config1"]
    h, V = CA.gen_integrals(8, 0)
    cim = CA.FCI(CA.Hamiltonian(h, V, 4, 4))
    ev, vec, info = CA.arpack_like(cim.get_mat().mvprod, cim.get_dim())
    assert info == 0 and abs(ev - g["energy"]) < 1e-10


def test_progress_callback_reports_convergence_tests():
    g = golden()["config1"]
    h, V = CA.gen_integrals(8, 0)
    cim = CA.FCI(CA.Hamiltonian(h, V, 4, 4))
    seen = []
    info = cim.solve(progress=lambda nmv, th, r: seen.append((nmv, th, r)))
    assert len(seen) >= 1 and seen[-1][0] <= info.n_matvec
    assert [s[0] for s in seen] == sorted(s[0] for s in seen)
    assert abs(seen[-1][1] - g["energy"]) < 1e-9


def test_fp32_values_with_fp64_polish():
    g = golden()["config1"]
    h, V = CA.gen_integrals(8, 0)
    cim = CA.FCI(CA.Hamiltonian(h, V, 4, 4), value_prec=CA.CF_F32)
    info = cim.solve(precision=CA.CF_F32)
    say("fp32+polish", info.as_dict())
    assert info.n_matvec_f32 > 0
    assert abs(cim.get_ci_energy() - g["energy"]) <= 1e-8


# ---- (d) RDMs ---------------------------------------------------------------
def _rdm_compare(L, one, two, o_one, o_two, tol=1e-9):
    assert np.max(np.abs(one - o_one)) <= tol
    assert np.max(np.abs(two - o_two)) <= tol


@pytest.mark.parametrize("case", ["fci_L6_3a3b", "fci_L8_4a4b_config1", "fci_beh2", "fci_L7_2a2b"])
def test_fci_rdms_match_oracle(case):
    kind, make = CASES[case]
    h, V, na, nb = make()
    L = h.shape[0]
    say("rdm", case)
    ham = CA.Hamiltonian(h, V, na, nb)
    cim = CA.FCI(ham)
    cim.solve()
    one, two = cim.construct_density()
    o_one, o_two = O.rdm_fci(L, na, nb, cim.get_eigvec())
    _rdm_compare(L, one, two, o_one, o_two)
    # invariants (CIDens.cpp:376, 398): tr gamma = N, sum Gamma(PQ,PQ) = N(N-1)/2
    N = na + nb
    tr = sum(cim.get_one_rdm(s, i, i) for s in range(2) for i in range(L))
    assert abs(tr - N) < 1e-11
    pairs = sum(cim.get_two_rdm(1, i, j, i, j) for i in range(L) for j in range(L))
    pairs += sum(cim.get_two_rdm(b, i, j, i, j) for b in (0, 2) for i in range(L) for j in range(i + 1, L))
    assert abs(pairs - N * (N - 1) / 2) < 1e-9


@pytest.mark.parametrize("case", ["doci_beh2", "doci_L10_5", "doci_redbcs_L8"])
def test_doci_rdms_match_oracle(case):
    kind, make = CASES[case]
    h, V, na, nb = make()
    L = h.shape[0]
    cim = CA.DOCI(CA.Hamiltonian(h, V, na, nb))
    cim.solve()
    one, two = cim.construct_density()
    o_one, o_two = O.rdm_doci(L, na, cim.get_eigvec())
    _rdm_compare(L, one, two, o_one, o_two)
    E = cim.get_ci_energy()
    assert abs(cim.get_dens_energy() - E) < 1e-10


# ---- benchmark-size workloads through size-independent checks ---------------
def test_config2_fci_L12_rows_bit_exact_and_counts():
    """FCI L=12 6a6b seed 1 (853,776 determinants): the whole CRS is built on
    the GPU; sampled alpha blocks are compared bit for bit with the oracle's
    construct_CI_matrix over the same alpha lines, and the totals with the
    closed-form counts (no element vanishes for these integrals)."""
    h, V = CA.gen_integrals(12, 1)
    say("config2 build")
    cim = CA.FCI(CA.Hamiltonian(h, V, 6, 6))
    n, nnz_full, nnz_up = cim.get_mat().info()
    assert n == 853776 and nnz_full == 1553018544 and nnz_up == 776936160
    db = 924
    for ia in (0, 401, 923):
        ref = O.build_crs("FCI", h, V, 6, 6, start=ia, end=ia + 1)
        row, col, val = cim.get_mat().export_upper(ia * db, (ia + 1) * db)
        assert np.array_equal(row, ref.row) and np.array_equal(col, ref.col) and np.array_equal(val, ref.data)


def test_config3_doci_L24_rows_bit_exact():
    h, V = CA.gen_integrals(24, 2)
    say("config3 build")
    cim = CA.DOCI(CA.Hamiltonian(h, V, 12, 12))
    n, nnz_full, nnz_up = cim.get_mat().info()
    assert n == 2704156 and nnz_full == 392102620 and nnz_up == 197403388
    for r0 in (0, 1352078, 2704156 - 300):
        ref = O.build_crs("DOCI", h, V, 12, 12, start=r0, end=r0 + 300)
        row, col, val = cim.get_mat().export_upper(r0, r0 + 300)
        assert np.array_equal(row, ref.row) and np.array_equal(col, ref.col) and np.array_equal(val, ref.data)


def _row_reference(h, V, na, nb, u_r, d_r, cand, x_of):
    """sum_j H_rj x_j over candidate determinants (u_j, d_j) with the oracle's
    get_ham_element, and sum_j |H_rj x_j| (the rounding scale)."""
    acc = scale = 0.0
    for uj, dj in cand:
        v = O.ham_element(h, V, na, nb, int(u_r), int(d_r), int(uj), int(dj)) * x_of(uj, dj)
        acc += v
        scale += abs(v)
    return acc, scale


def test_config2_default_product_sampled_rows_match_oracle_elements():
    """FCI L=12 6a6b seed 1 at full size: the default FP64 product (the
    position-major ELL kernel) and the FP32-value one on sampled rows against
    the oracle's get_ham_element summed over every determinant within a double
    excitation (found by popcount over the whole space, independent of the
    GPU's pattern)."""
    L, na, nb = 12, 6, 6
    h, V = CA.gen_integrals(L, 1)
    cim = CA.FCI(CA.Hamiltonian(h, V, na, nb), value_prec=CA.CF_F32)
    m = cim.get_mat()
    assert m.spmv_bytes(CA.CF_F64)[0] == 4
    strs = O.enumerate_strings(L, na)
    d = len(strs)
    rank = {int(s): k for k, s in enumerate(strs)}
    x = np.random.default_rng(12).standard_normal(cim.get_dim())
    y64 = m.mvprod(x, CA.CF_F64)
    y32 = m.mvprod(x, CA.CF_F32)
    say("config2 product rows")
    for r in (0, 123457, 500000, 853775):
        ia, ib = divmod(r, d)
        ca = [strs[j] for j in _connected(strs, strs[ia], 2)]
        cb = [strs[j] for j in _connected(strs, strs[ib], 2)]
        cand = [(ua, ub) for ua in ca for ub in cb
                if bin(int(strs[ia]) ^ int(ua)).count("1") + bin(int(strs[ib]) ^ int(ub)).count("1") <= 4]
        acc, scale = _row_reference(h, V, na, nb, strs[ia], strs[ib], cand,
                                    lambda u, v: x[rank[int(u)] * d + rank[int(v)]])
        assert abs(y64[r] - acc) <= 1e-13 * scale, (r, y64[r], acc)
        assert abs(y32[r] - acc) <= 1e-6 * scale, (r, y32[r], acc)


def test_config3_default_product_sampled_rows_match_oracle_elements():
    """DOCI L=24 12 pairs seed 2 at full size: the default product (the
    block-staged kernel, which stores no columns) on sampled rows against the
    oracle's get_ham_element between the closed-shell determinants (u, u) and
    (v, v) -- DOCI's diagonal and pair hop V(j,j,k,k) (CIMethod.cpp:611-660) --
    summed over every pair string one hop away (popcount over all 2,704,156
    strings)."""
    L, npair = 24, 12
    h, V = CA.gen_integrals(L, 2)
    cim = CA.DOCI(CA.Hamiltonian(h, V, npair, npair))
    m = cim.get_mat()
    assert m.spmv_bytes(CA.CF_F64)[0] == 6
    strs = O.enumerate_strings(L, npair)
    x = np.random.default_rng(24).standard_normal(cim.get_dim())
    y = m.mvprod(x, CA.CF_F64)
    say("config3 product rows")
    for r in (0, 999999, 1352078, 2704155):
        u = strs[r]
        js = np.nonzero(np.bitwise_count(strs ^ u) <= 2)[0]
        assert len(js) == 1 + npair * (L - npair)
        acc, scale = _row_reference(h, V, npair, npair, u, u, [(strs[j], strs[j]) for j in js],
                                    lambda a, b: x[int(np.searchsorted(strs, a))])
        assert abs(y[r] - acc) <= 1e-13 * scale, (r, y[r], acc)


# ---- (e) integral rotation on FP64 MFMA ---------------------------------------
@pytest.mark.parametrize("L", [4, 7, 12, 20])
def test_rotate_integrals_match_oracle(L):
    say("rotate", L)
    B = 3
    hs, Vs, Us = [], [], []
    for b in range(B):
        h, V = CA.gen_integrals(L, 100 + b)
        hs.append(h)
        Vs.append(V)
        Us.append(CA.expm_skew(CA.gen_rotation(L, 1000 + b, 0.05 if b else 0.8)))
    ho, Vo = CA.rotate_integrals(np.array(hs), np.array(Vs), np.array(Us))
    for b in range(B):
        rh, rV = O.rotate_integrals(Us[b], hs[b], Vs[b])
        assert np.max(np.abs(ho[b] - rh)) < 1e-13
        assert np.max(np.abs(Vo[b] - rV)) < 1e-13 * max(1.0, np.max(np.abs(rV)))


def test_rotation_round_trip_and_fci_invariance():
    """TestTransformations.cpp:120 (U then U^-1 restores the integrals) and
    E_FCI(U . ints) = E_FCI(ints) through the GPU build and solve."""
    g = golden()["config1"]
    h, V = CA.gen_integrals(8, 0)
    ham = CA.Hamiltonian(h, V, 4, 4)
    U = CA.expm_skew(CA.gen_rotation(8, 7, 0.3))
    hr = ham.rotate(U)
    back = hr.rotate(U.T)
    assert np.max(np.abs(back.h - h)) < 1e-13 and np.max(np.abs(back.V - V)) < 1e-13
    cim = CA.FCI(hr)
    cim.solve()
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-10


# ---- (c) staged FCI SpMV ------------------------------------------------------
@pytest.mark.parametrize("L,n,W", [(8, 4, 16), (10, 5, 8), (11, 5, 16), (12, 6, 16), (12, 6, 32)])
def test_staged_fci_spmv_matches_row_gather(L, n, W):
    say("staged spmv", L, n, W)
    h, V = CA.gen_integrals(L, 1)
    cim = CA.FCI(CA.Hamiltonian(h, V, n, n))
    m = cim.get_mat()
    x = np.random.default_rng(2).standard_normal(cim.get_dim())
    m.set_spmv_kernel(1)
    y1 = m.mvprod(x)
    m.set_spmv_kernel(2, W)
    y2 = m.mvprod(x)
    assert np.max(np.abs(y2 - y1)) <= 1e-13 * np.max(np.abs(y1))


def test_staged_spmv_falls_back_on_dropped_entries():
    """D2h fixture: symmetry zeros are dropped, rows are shorter than the
    template, the staged kernel must take its plain-gather path."""
    h, V, na, nb, ec = fixture("beh2_sto3g.npz")
    cim = CA.FCI(CA.Hamiltonian(h, V, na, nb, ec))
    m = cim.get_mat()
    x = np.random.default_rng(3).standard_normal(cim.get_dim())
    m.set_spmv_kernel(1)
    y1 = m.mvprod(x)
    m.set_spmv_kernel(2, 8)
    y2 = m.mvprod(x)
    assert np.max(np.abs(y2 - y1)) <= 1e-13 * np.max(np.abs(y1))
    cim.solve()
    assert abs(cim.get_ci_energy() - golden()["beh2"]["fci_energy"]) < 1e-10


@pytest.mark.parametrize("case", ["fci_L8_4a4b_config1", "fci_beh2", "doci_beh2", "doci_L10_5", "fci_L11_5a5b",
                                  "doci_L20_10"])
def test_vectorised_spmv_matches_scalar_gather(case):
    """Kernel 3 (vectorised column/value streams, unaligned row heads and
    tails) and the automatic choice against the scalar row gather, FP64 and
    FP32 values; D2h rows (dropped entries) have lengths that are not
    multiples of 4."""
    say("v4 spmv", case)
    if case in CASES:
        kind, make = CASES[case]
        h, V, na, nb = make()
    elif case == "fci_L11_5a5b":
        (h, V), kind, na, nb = CA.gen_integrals(11, 1), "FCI", 5, 5
    else:
        (h, V), kind, na, nb = CA.gen_integrals(20, 1), "DOCI", 10, 10
    cls = CA.FCI if kind == "FCI" else CA.DOCI
    cim = cls(CA.Hamiltonian(h, V, na, nb), value_prec=CA.CF_F32)
    m = cim.get_mat()
    x = np.random.default_rng(5).standard_normal(cim.get_dim())
    for prec in (CA.CF_F64, CA.CF_F32):
        m.set_spmv_kernel(1)
        y1 = m.mvprod(x, prec)
        for kern in (3, 0):
            m.set_spmv_kernel(kern)
            y2 = m.mvprod(x, prec)
            assert np.max(np.abs(y2 - y1)) <= 1e-13 * np.max(np.abs(y1)), (case, prec, kern)


@pytest.mark.parametrize("case,nt", [("fci_L8_4a4b_config1", 512), ("fci_beh2", 256), ("fci_L7_3a2b", 512),
                                     ("fci_L11_5a5b", 512), ("fci_L11_5a5b", 1024), ("fci_L12_6a6b", 256)])
def test_ell_spmv_matches_scalar_gather(case, nt):
    """Kernel 4 (position-major complete rows, LDS-staged x blocks): zeros at
    the dropped D2h positions, nalpha != nbeta, several workgroup sizes."""
    say("ell spmv", case, nt)
    if case in CASES:
        _, make = CASES[case]
        h, V, na, nb = make()
    else:
        L, na, nb = {"fci_L7_3a2b": (7, 3, 2), "fci_L11_5a5b": (11, 5, 5), "fci_L12_6a6b": (12, 6, 6)}[case]
        h, V = CA.gen_integrals(L, 1)
    cim = CA.FCI(CA.Hamiltonian(h, V, na, nb), value_prec=CA.CF_F32)
    m = cim.get_mat()
    x = np.random.default_rng(6).standard_normal(cim.get_dim())
    for prec in (CA.CF_F64, CA.CF_F32):
        m.set_spmv_kernel(1)
        y1 = m.mvprod(x, prec)
        m.set_spmv_kernel(4, nt)
        y2 = m.mvprod(x, prec)
        assert np.max(np.abs(y2 - y1)) <= 1e-13 * np.max(np.abs(y1)), (case, prec)
    if case == "fci_beh2":  # the solve on kernel 4 (fixture Econst added)
        cim.solve()
        econst = fixture("beh2_sto3g.npz")[4]
        assert abs(cim.get_ci_energy() + econst - golden()["beh2"]["fci_energy"]) < 1e-10


DOCI_HOP_CASES = {
    "doci_L9_3": (9, 3, 4),
    "doci_L6_1": (6, 1, 5),
    "doci_L7_6": (7, 6, 6),
    "doci_L14_7_dropped": (14, 7, 7),
    "doci_L20_10": (20, 10, 8),
}


@pytest.mark.parametrize("case", ["doci_beh2", "doci_L10_5", "doci_redbcs_L8"] + list(DOCI_HOP_CASES))
def test_doci_hop_spmv_matches_scalar_gather(case):
    """Kernel 5 (position-major pair-hop columns, pair elements from an LDS
    table) and kernel 6 (block-staged, no columns; several low/high splits)
    against the scalar row gather of the stored CRS, FP64 and FP32 values,
    1-8 lanes per row (more lanes than pairs included): BeH2 (D2h zeros
    dropped by the build), one pair, one hole, and a threshold that drops
    half of the pair elements."""
    say("doci hop spmv", case)
    zero = 1e-14
    if case in CASES:
        _, make = CASES[case]
        h, V, na, nb = make()
    else:
        L, na, seed = DOCI_HOP_CASES[case]
        nb = na
        h, V = CA.gen_integrals(L, seed)
        if case.endswith("dropped"):
            m0 = CA.DOCI(CA.Hamiltonian(h, V, na, nb)).get_mat()
            row, col, val = m0.export_full(0, 64)
            off = np.abs(val[col != np.repeat(np.arange(64), np.diff(row.astype(np.int64)))])
            zero = float(np.median(off))
            del m0
    cim = CA.DOCI(CA.Hamiltonian(h, V, na, nb), value_prec=CA.CF_F32, zero=zero)
    m = cim.get_mat()
    n = cim.get_dim()
    if case.endswith("dropped"):
        assert m.info()[1] < n * (1 + na * (len(h) - na)) * 0.75  # elements really dropped
    # automatic choice: block-staged from 60,000 rows, hop columns from
    # 40,000 (FP64 values), the CRS row gather below
    k64, k32 = m.spmv_bytes(CA.CF_F64)[0], m.spmv_bytes(CA.CF_F32)[0]
    assert k64 == (6 if n >= 60000 else 5 if n >= 40000 else k64) and (n >= 40000 or k64 < 4)
    assert k32 == (6 if n >= 60000 else k32) and (n >= 60000 or k32 < 4)
    x = np.random.default_rng(9).standard_normal(n)
    for prec in (CA.CF_F64, CA.CF_F32):
        m.set_spmv_kernel(1)
        y1 = m.mvprod(x, prec)
        m.set_spmv_kernel(5)  # builds the hop columns on request
        y2 = m.mvprod(x, prec)
        assert np.max(np.abs(y2 - y1)) <= 1e-13 * np.max(np.abs(y1)), (case, prec)
        for nt, sub in ((64, 1), (256, 2), (1024, 4), (128, 8), (256, 8)):
            m.set_spmv_shape(prec, 5, nt, sub)
            y2 = m.mvprod(x, prec)
            assert np.max(np.abs(y2 - y1)) <= 1e-13 * np.max(np.abs(y1)), (case, prec, nt, sub)
        for w in sorted({max(1, len(h) // 2 - 1), len(h) // 2, min(12, len(h) - 1)}):
            m.set_spmv_shape(prec, 6, w, 0)  # block-staged, w low orbitals
            y3 = m.mvprod(x, prec)
            assert np.max(np.abs(y3 - y1)) <= 1e-13 * np.max(np.abs(y1)), (case, prec, "blk", w)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 5, 100, 1)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 5, 256, 3)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 6, 13, 0)


def test_doci_hop_kernel_rejected_for_fci():
    h, V = CA.gen_integrals(6, 7)
    m = CA.FCI(CA.Hamiltonian(h, V, 3, 3)).get_mat()
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_kernel(5)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 5, 256, 1)


@pytest.mark.parametrize("L,na,nb", [(10, 5, 5), (7, 3, 2)])
def test_autotune_shapes_agree_and_fastest_is_installed(L, na, nb):
    """Every launch shape the autotuner times gives the scalar gather's product
    (within rounding), the fastest becomes the automatic choice, and shapes the
    kernels cannot run are rejected."""
    say("autotune", L, na, nb)
    h, V = CA.gen_integrals(L, 1)
    cim = CA.FCI(CA.Hamiltonian(h, V, na, nb), value_prec=CA.CF_F32)
    m = cim.get_mat()
    m.set_spmv_kernel(4, 0)  # builds the position-major copy
    x = np.random.default_rng(3).standard_normal(cim.get_dim())
    for prec in (CA.CF_F64, CA.CF_F32):
        m.set_spmv_kernel(1)
        y0 = m.mvprod(x, prec)
        tab = m.autotune(prec, reps=2)
        kinds = {k for k, _, _, _ in tab}
        # the sliced copy (kernel 7) is made at build for FCI spaces of >= 8,192 rows
        assert kinds == ({1, 3, 4, 7} if cim.get_dim() >= 8192 else {1, 3, 4})
        best = min(tab, key=lambda t: t[3])
        assert m.spmv_bytes(prec)[0] == best[0]
        yb = m.mvprod(x, prec)
        assert np.max(np.abs(yb - y0)) <= 1e-13 * np.max(np.abs(y0))
        for k, p1, p2, ms in tab:
            assert ms > 0
            m.set_spmv_shape(prec, k, p1, p2)
            y = m.mvprod(x, prec)
            assert np.max(np.abs(y - y0)) <= 1e-13 * np.max(np.abs(y0)), (prec, k, p1, p2)
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 4, 100, 1)  # not a multiple of 64
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 3, 64, 1)  # lanes per row
    with pytest.raises(CA.CIFlowError):
        m.set_spmv_shape(CA.CF_F64, 4, 256, 3)  # strings per thread
    if L == 7:  # d_beta = 21 is odd
        with pytest.raises(CA.CIFlowError):
            m.set_spmv_shape(CA.CF_F64, 4, 256, 2)


def test_config2_energy_matches_oracle():
    """FCI L=12 6a6b seed 1: GPU solve vs the oracle's full CPU solve (golden)."""
    g = golden()["config2"]
    h, V = CA.gen_integrals(12, 1)
    cim = CA.FCI(CA.Hamiltonian(h, V, 6, 6), value_prec=CA.CF_F32)
    info = cim.solve()
    say("config2 f64", info.as_dict())
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-10
    info = cim.solve(precision=CA.CF_F32)
    say("config2 f32+polish", info.as_dict())
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-8


def test_config3_doci_L24_energy_matches_oracle():
    """DOCI L=24 12 pairs seed 2 (2,704,156 determinants): GPU build + solve vs
    the oracle's full CPU solve (golden), FP64, FP32+polish and the served
    energy stop."""
    g = golden()["config3"]
    h, V = CA.gen_integrals(24, 2)
    cim = CA.DOCI(CA.Hamiltonian(h, V, 12, 12), value_prec=CA.CF_F32)
    n, nnz_full, nnz_upper = cim.get_mat().info()
    assert n == 2704156 and nnz_upper == g["nnz_upper"] and nnz_full == 2 * nnz_upper - n
    info = cim.solve()
    say("config3 f64", info.as_dict())
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-10
    info = cim.solve(precision=CA.CF_F32)
    say("config3 f32+polish", info.as_dict())
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-8
    info = cim.solve(precision=CA.CF_F32, tol_energy=1e-9, check_every=4)
    say("config3 f32+energy stop", info.as_dict())
    assert abs(cim.get_ci_energy() - g["energy"]) < 1e-8


# ---- (c) on-the-fly sigma (no stored matrix) -----------------------------------
@pytest.mark.parametrize("L,na,nb", [(6, 3, 3), (8, 4, 4), (10, 5, 5), (7, 3, 2)])
def test_direct_sigma_matches_crs_spmv(L, na, nb):
    say("sigma", L, na, nb)
    h, V = CA.gen_integrals(L, 3)
    ham = CA.Hamiltonian(h, V, na, nb)
    crs = CA.FCI(ham)
    direct = CA.FCI_Direct(ham, value_prec=CA.CF_F32)
    x = np.random.default_rng(4).standard_normal(crs.get_dim())
    y0 = crs.get_mat().mvprod(x)
    y1 = direct.mvprod(x)
    assert np.max(np.abs(y1 - y0)) <= 1e-12 * np.max(np.abs(y0))
    y2 = direct.mvprod(x, precision=CA.CF_F32)
    assert np.max(np.abs(y2 - y0)) <= 3e-6 * np.max(np.abs(y0))


@pytest.mark.parametrize("L,na,nb", [(6, 3, 3), (8, 4, 4), (9, 4, 4), (11, 5, 5), (12, 6, 6), (7, 3, 2)])
def test_sigma_alpha_beta_kernels_agree(L, na, nb):
    """Kernel 3 (packed offsets, the default) against kernel 2 and the stored
    CRS SpMV: FP64 to 1e-12 relative, FP32 operands to the FP32 rounding."""
    say("sigma kernels", L, na, nb)
    h, V = CA.gen_integrals(L, 4)
    ham = CA.Hamiltonian(h, V, na, nb)
    direct = CA.FCI_Direct(ham, value_prec=CA.CF_F32)
    x = np.random.default_rng(6).standard_normal(direct.get_dim())
    ys = {}
    for kern in (2, 3):
        direct.set_kernel(kern)
        ys[kern] = (direct.mvprod(x), direct.mvprod(x, precision=CA.CF_F32))
    scale = np.max(np.abs(ys[2][0]))
    assert np.max(np.abs(ys[3][0] - ys[2][0])) <= 1e-12 * scale
    assert np.max(np.abs(ys[3][1] - ys[2][0])) <= 3e-6 * scale
    if L <= 9:
        y0 = CA.FCI(ham).get_mat().mvprod(x)
        assert np.max(np.abs(ys[3][0] - y0)) <= 1e-12 * scale


def test_direct_solve_and_rdm1_match_crs_path():
    h, V = CA.gen_integrals(10, 5)
    ham = CA.Hamiltonian(h, V, 5, 5)
    crs = CA.FCI(ham)
    crs.solve()
    one_ref, _ = crs.construct_density(trdm=False)
    direct = CA.FCI_Direct(ham, value_prec=CA.CF_F32)
    info = direct.solve()
    say("direct f64", info.as_dict())
    assert abs(direct.get_ci_energy() - crs.get_ci_energy()) < 1e-10
    one, _ = direct.construct_density()
    assert np.max(np.abs(one - one_ref)) < 1e-10
    info = direct.solve(precision=CA.CF_F32)
    say("direct f32+polish", info.as_dict())
    assert abs(direct.get_ci_energy() - crs.get_ci_energy()) < 1e-8


def _connected(strs, s, nex):
    return [j for j, t in enumerate(strs) if bin(int(s) ^ int(t)).count("1") <= 2 * nex]


def test_direct_sigma_L16_sampled_rows_match_oracle_elements():
    """FCI L=16 8a8b (165,636,900 determinants): sigma rows against the
    oracle's get_ham_element (CIMethod.cpp:459-487) summed over every
    determinant within a double excitation of the row."""
    L, n = 16, 8
    h, V = CA.gen_integrals(L, 3)
    ham = CA.Hamiltonian(h, V, n, n)
    say("config5 plan")
    direct = CA.FCI_Direct(ham)
    d = direct.da
    strs = O.enumerate_strings(L, n)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(direct.get_dim())
    y = direct.mvprod(x)
    for r in (0, 1234567, direct.get_dim() - 1, 98765432):
        ia, ib = divmod(r, d)
        acc = 0.0
        ca = _connected(strs, strs[ia], 2)
        cb = _connected(strs, strs[ib], 2)
        for ja in ca:
            ea = bin(int(strs[ia]) ^ int(strs[ja])).count("1") // 2
            for jb in cb:
                eb = bin(int(strs[ib]) ^ int(strs[jb])).count("1") // 2
                if ea + eb > 2:
                    continue
                acc += O.ham_element(h, V, n, n, int(strs[ia]), int(strs[ib]), int(strs[ja]), int(strs[jb])) * \
                    x[ja * d + jb]
        assert abs(y[r] - acc) <= 1e-11 * max(1.0, abs(acc)), (r, y[r], acc)
