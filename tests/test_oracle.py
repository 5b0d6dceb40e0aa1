"""CPU tests that pin the oracle (oracle/ciflow_oracle.cpp) before it is used
as the checker of the GPU library.

Anchors: brute force over bitstrings, brute-force second quantisation
(oracle/bruteforce.py), numpy eigh, the reference's known-answer tests
(TestDOCI.cpp:31, TestDensity.cpp:57, TestRedBCS.cpp:44-45) on the
reference's own integral dumps (converted to tests/golden/*.npz), and the
committed golden digests.
"""
import hashlib
import itertools
import json
import math
import os

import numpy as np
import pytest

from oracle import bruteforce as BF
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def golden():
    with open(os.path.join(GOLD, "golden.json")) as f:
        return json.load(f)


def sha(*arrays):
    m = hashlib.sha256()
    for a in arrays:
        m.update(np.ascontiguousarray(a).tobytes())
    return m.hexdigest()


# --- a1/a2: enumeration and ranking -----------------------------------------
@pytest.mark.parametrize("L,n", [(1, 0), (1, 1), (5, 2), (6, 3), (8, 4), (10, 3), (12, 6)])
def test_gosper_order_is_increasing_integers_and_rank_is_position(L, n):
    strs = O.enumerate_strings(L, n)
    brute = sorted(sum(1 << o for o in occ) for occ in itertools.combinations(range(L), n))
    assert [int(s) for s in strs] == brute
    for i, s in enumerate(brute):
        assert O.rank(L, n, s) == i
        # closed form of the vertex-weight rank: sum_m C(k_m, m)
        ks = [k for k in range(L) if (s >> k) & 1]
        assert sum(math.comb(k, m + 1) for m, k in enumerate(ks)) == i


# --- integral generator -------------------------------------------------------
@pytest.mark.parametrize("key", ["8_0", "12_1", "4_5", "6_7"])
def test_generator_matches_committed_digest(key):
    L, seed = map(int, key.split("_"))
    h, V = O.gen_integrals(L, seed)
    assert sha(h, V) == golden()["integrals"][key]


def test_generator_closed_form_properties():
    L = 6
    h, V = O.gen_integrals(L, 7)
    assert np.array_equal(h, h.T)
    assert np.allclose(np.diag(h), -2 + 3 * np.arange(L) / (L - 1), rtol=0, atol=0)
    # physics V[i,k,j,l] = (ij|kl): exact 8-fold symmetry of the chemist tensor
    chem = V.transpose(0, 2, 1, 3)
    for perm in [(1, 0, 2, 3), (0, 1, 3, 2), (2, 3, 0, 1), (3, 2, 1, 0)]:
        assert np.array_equal(chem, chem.transpose(perm))
    # positive semidefinite as a (ij),(kl) matrix
    assert np.linalg.eigvalsh(chem.reshape(L * L, L * L)).min() > -1e-12
    # off-diagonal h ~ N(0, 0.05^2): loose moment check over a bigger draw
    h2, _ = O.gen_integrals(24, 2)
    off = h2[np.triu_indices(24, 1)]
    assert abs(off.std() - 0.05) < 0.01 and abs(off.mean()) < 0.01


# --- a3/a4/a6/a7: Slater-Condon elements and RDM conventions vs brute force --
@pytest.mark.parametrize("L,na,nb,seed", [(4, 2, 2, 5), (5, 2, 2, 7), (5, 3, 3, 11), (4, 1, 1, 3)])
def test_fci_matrix_and_rdms_match_second_quantisation(L, na, nb, seed):
    h, V = O.gen_integrals(L, seed)
    m = O.build_crs("FCI", h, V, na, nb)
    H = O.dense_from_upper(m)
    Hb = BF.hamiltonian(h, V, na, nb)
    assert np.abs(H - Hb).max() < 1e-13
    w, U = np.linalg.eigh(H)
    c = U[:, 0]
    one, two = O.rdm_fci(L, na, nb, c)
    g, G = BF.rdms(c, L, na, nb)
    for s, a, b in itertools.product(range(2), range(L), range(L)):
        assert abs(O.get_one_rdm(L, one, s, a, b) - g[s, a, b]) < 1e-13
    for blk, i, j, k, l in itertools.product(range(3), *(range(L),) * 4):
        assert abs(O.get_two_rdm(L, two, blk, i, j, k, l) - G[blk, i, j, k, l]) < 1e-13
    assert abs(O.dens_energy(h, V, 0.0, one, two) - w[0]) < 1e-12


def test_doci_matrix_and_rdms_match_seniority_zero_block():
    L, n = 6, 3
    h, V = O.gen_integrals(L, 3)
    m = O.build_crs("DOCI", h, V, n, n)
    H = O.dense_from_upper(m)
    dets = [(a, a) for a in BF.strings(L, n)]
    Hb = BF.hamiltonian(h, V, n, n, dets=dets)
    assert np.abs(H - Hb).max() < 1e-13
    w, U = np.linalg.eigh(H)
    one, two = O.rdm_doci(L, n, U[:, 0])
    g, G = BF.rdms(U[:, 0], L, n, n, dets=dets)
    for blk, i, j, k, l in itertools.product(range(3), *(range(L),) * 4):
        assert abs(O.get_two_rdm(L, two, blk, i, j, k, l) - G[blk, i, j, k, l]) < 1e-13
    for s, a, b in itertools.product(range(2), range(L), range(L)):
        assert abs(O.get_one_rdm(L, one, s, a, b) - g[s, a, b]) < 1e-13


# --- reference known-answer tests ------------------------------------------
def load_fixture(name):
    d = np.load(os.path.join(GOLD, name))
    return d["h"], d["V"], int(d["L"]), int(d["nalpha"]), int(d["nbeta"]), float(d["econst"])


def test_doci_beh2_kat():
    """TestDOCI.cpp:31: E = -15.225362047972723, BOOST_CHECK_CLOSE 1e-10 % (1e-12 rel)."""
    h, V, L, na, nb, ec = load_fixture("beh2_sto3g.npz")
    m = O.build_crs("DOCI", h, V, na, na)
    E, _, info = O.lanczos(O.full_from_upper(m))
    assert info["status"] == 0
    assert abs((E + ec) - (-15.225362047972723)) <= 1e-12 * 15.23


def test_fci_beh2_density_energy_identity():
    """TestDensity.cpp:57: E_dens == E_FCI (1e-12 rel)."""
    h, V, L, na, nb, ec = load_fixture("beh2_sto3g.npz")
    m = O.build_crs("FCI", h, V, na, nb)
    E, c, _ = O.lanczos(O.full_from_upper(m))
    one, two = O.rdm_fci(L, na, nb, c)
    assert abs(O.dens_energy(h, V, ec, one, two) - (E + ec)) <= 1e-12 * abs(E + ec)
    assert abs(E + ec - golden()["beh2"]["fci_energy"]) < 1e-10


def redbcs_ham(L, eps, g):
    """RedBCS model (ModHam.cpp:174-205, RedBCS.cpp:38-121): T_ii = eps_i,
    getVmat(i,i,j,j) = g, and _modham zeroes every V with i != j
    (Hamiltonian.cpp:311-318)."""
    h = np.diag(np.asarray(eps, dtype=float))
    V = np.zeros((L, L, L, L))
    for i in range(L):
        for j in range(L):
            V[i, i, j, j] = g
    return h, V


def test_redbcs_kat():
    """TestRedBCS.cpp:44-45: E_red = -2.5978142225, E_var = -1.1247132522 (1e-9 rel)."""
    h0, V0, L, na, nb, ec = load_fixture("psi0_631g.npz")
    eps = np.diag(h0).copy()
    perm = np.argsort(eps, kind="stable")
    h, V = redbcs_ham(L, np.sort(eps), 0.119034351059)
    m = O.build_crs("DOCI", h, V, na, na)
    H = O.dense_from_upper(m)
    w, U = np.linalg.eigh(H)
    e_red = w[0]
    assert abs(e_red - (-2.5978142225)) <= 1e-9 * 2.6
    c = U[:, 0]
    one, two = O.rdm_doci(L, na, c)
    # RedBCS::get_energy(ham), RedBCS.cpp:150-175
    e = 0.0
    for i in range(L):
        e += 2.0 * h0[perm[i], perm[i]] * O.get_one_rdm(L, one, 0, i, i)
    for i in range(L):
        for j in range(L):
            if i != j:
                v2 = V0[perm[i], perm[j], perm[i], perm[j]]
                v3 = V0[perm[i], perm[j], perm[j], perm[i]]
                v1 = V0[perm[i], perm[i], perm[j], perm[j]]
                e += (2 * v2 - v3) * O.get_two_rdm(L, two, 1, i, j, i, j)
                e += v1 * O.get_two_rdm(L, two, 1, i, i, j, j)
        e += V0[perm[i], perm[i], perm[i], perm[i]] * O.get_two_rdm(L, two, 1, i, i, i, i)
    e += ec
    assert abs(e - (-1.1247132522)) <= 1e-9 * 1.13


# --- a6-a9: CRS contract, golden digest at the benchmark's config 1 ----------
def test_config1_crs_digest_and_energy():
    g = golden()["config1"]
    h, V = O.gen_integrals(8, 0)
    m = O.build_crs("FCI", h, V, 4, 4)
    assert m.nnz == g["nnz_upper"] == 886900
    assert sha(m.row.astype(np.uint64), m.col.astype(np.uint32)) == g["row_col_sha256"]
    assert sha(np.signbit(m.data).astype(np.uint8)) == g["sign_sha256"]
    E, _, info = O.lanczos(O.full_from_upper(m))
    assert info["status"] == 0
    assert abs(E - g["energy"]) < 1e-10


def test_build_is_independent_of_thread_partition():
    h, V = O.gen_integrals(6, 7)
    a = O.build_crs("FCI", h, V, 3, 3, threads=1)
    b = O.build_crs("FCI", h, V, 3, 3, threads=7)
    assert np.array_equal(a.row, b.row) and np.array_equal(a.col, b.col) and np.array_equal(a.data, b.data)


def test_drop_rule_on_symmetry_zeros():
    """D2h fixture: symmetry-forbidden elements are exact zeros and are not
    stored (PushToRow |v| > 1e-14, SparseMatrix_CRS.cpp:242)."""
    h, V, L, na, nb, ec = load_fixture("beh2_sto3g.npz")
    m = O.build_crs("FCI", h, V, na, nb)
    assert np.all(np.abs(m.data) > 1e-14)
    full_pattern = 1 + 2 * 12 + 144 + 2 * 36  # per-row upper bound of connected dets at L=7, 3a3b
    assert m.nnz < m.n * full_pattern / 2 and m.nnz == golden()["beh2"]["fci_crs"]["nnz_upper"]


def test_mvprod_variants_agree():
    h, V = O.gen_integrals(6, 7)
    m = O.build_crs("FCI", h, V, 3, 3)
    f = O.full_from_upper(m)
    H = O.dense_from_upper(m)
    x = np.random.default_rng(0).standard_normal(m.n)
    assert np.allclose(O.mvprod_sym(m, x), H @ x, rtol=0, atol=1e-12)
    assert np.allclose(O.mvprod_full(f, x), H @ x, rtol=0, atol=1e-12)


@pytest.mark.parametrize("L,na,seed", [(6, 3, 7), (7, 3, 2)])
def test_lanczos_matches_eigh(L, na, seed):
    h, V = O.gen_integrals(L, seed)
    m = O.build_crs("FCI", h, V, na, na)
    w, U = np.linalg.eigh(O.dense_from_upper(m))
    E, x, info = O.lanczos(O.full_from_upper(m))
    assert info["status"] == 0
    assert abs(E - w[0]) < 1e-11
    assert abs(abs(x @ U[:, 0]) - 1) < 1e-9


# --- a15: integral rotation ---------------------------------------------------
def test_rotation_oracle_matches_einsum_and_round_trips():
    import scipy.linalg as sl
    L = 6
    h, V = O.gen_integrals(L, 7)
    rng = np.random.default_rng(1)
    K = rng.normal(0, 0.3, (L, L))
    K = K - K.T
    U = sl.expm(K)
    ho, Vo = O.rotate_integrals(U, h, V)
    assert np.max(np.abs(ho - U @ h @ U.T)) < 1e-13
    ref = np.einsum("ai,bj,ck,dl,ijkl->abcd", U, U, U, U, V, optimize=True)
    assert np.max(np.abs(Vo - ref)) < 1e-13
    # TestTransformations.cpp:120: U then U^-1 gives back the integrals
    h2, V2 = O.rotate_integrals(U.T, ho, Vo)
    assert np.max(np.abs(h2 - h)) < 1e-13 and np.max(np.abs(V2 - V)) < 1e-13
    # FCI energy is invariant under orbital rotation
    e0 = np.linalg.eigvalsh(O.dense_from_upper(O.build_crs("FCI", h, V, 3, 3)))[0]
    e1 = np.linalg.eigvalsh(O.dense_from_upper(O.build_crs("FCI", ho, Vo, 3, 3)))[0]
    assert abs(e0 - e1) < 1e-11


# --- selected CI from determinant lists (FCI_File, Permutator_File) ----------
def random_det_list(L, na, nb, count, seed):
    dets = BF.basis(L, na, nb)
    rng = np.random.default_rng(seed)
    pick = rng.permutation(len(dets))[:count]  # a subset in shuffled order
    up = np.array([dets[i][0] for i in pick], dtype=np.uint64)
    down = np.array([dets[i][1] for i in pick], dtype=np.uint64)
    return up, down


@pytest.mark.parametrize("L,na,nb,count,seed", [(5, 2, 2, 60, 1), (6, 3, 2, 90, 2), (6, 3, 3, 400, 3)])
def test_det_list_matrix_matches_second_quantisation(L, na, nb, count, seed):
    """FCI_File::construct_CI_matrix on an arbitrary ordered subset: the dense
    matrix equals <I|H|J> from explicit creation/annihilation operators."""
    h, V = O.gen_integrals(L, seed + 10)
    up, down = random_det_list(L, na, nb, count, seed)
    m = O.build_crs_dets(h, V, na, nb, up, down)
    assert m.n == count
    H = O.dense_from_upper(m)
    Hb = BF.hamiltonian(h, V, na, nb, dets=[(int(a), int(b)) for a, b in zip(up, down)])
    assert np.abs(H - Hb).max() < 1e-13


def test_det_file_be_kat():
    """TestDetFile.cpp:32: Be cc-pVDZ over the 100-determinant GKCI list,
    E = -14.5951673370889 (BOOST_CHECK_CLOSE 1e-10 %, i.e. 1e-12 relative)."""
    h, V, L, na, nb, ec = load_fixture("be_ccpvdz.npz")
    up, down = O.read_dets(os.path.join(GOLD, "GKCI_Be_cntSp_cc-pVDZ_0100.txt"))
    assert len(up) == 100
    m = O.build_crs_dets(h, V, na, nb, up, down)
    E, _, info = O.lanczos(O.full_from_upper(m))
    assert info["status"] == 0
    assert abs((E + ec) - (-14.5951673370889)) <= 1e-12 * 14.6


def test_det_file_full_beh2_list_is_the_permuted_fci_matrix():
    """The BeH2 file lists all 35 x 35 determinants in its own order: its matrix is
    the FCI matrix permuted (same pattern, elements to 1e-12) and its ground
    state is the FCI energy."""
    h, V, L, na, nb, ec = load_fixture("beh2_sto3g.npz")
    up, down = O.read_dets(os.path.join(GOLD, "fci_beh2_sto-3gdeterminants.dat"))
    m = O.build_crs_dets(h, V, na, nb, up, down)
    Hd = O.dense_from_upper(m)
    A = O.enumerate_strings(L, na)
    B = O.enumerate_strings(L, nb)
    pos = np.array([O.rank(L, na, int(a)) * len(B) + O.rank(L, nb, int(b)) for a, b in zip(up, down)])
    assert sorted(pos.tolist()) == list(range(len(A) * len(B)))
    Hf = O.dense_from_upper(O.build_crs("FCI", h, V, na, nb))[np.ix_(pos, pos)]
    assert np.array_equal(Hd != 0, Hf != 0)
    assert np.abs(Hd - Hf).max() <= 1e-12 * np.abs(Hf).max()
    E, _, _ = O.lanczos(O.full_from_upper(m))
    assert abs((E + ec) - golden()["beh2"]["fci_energy"]) < 1e-10


@pytest.mark.parametrize("L,npair,g", [(4, 2, -0.5), (6, 3, -0.5), (8, 4, -0.5), (10, 5, -0.5), (12, 6, -0.5),
                                       (12, 6, 0.3), (10, 3, -1.0), (12, 4, -0.2), (9, 4, -0.5), (11, 2, 0.7)])
def test_richardson_gaudin_matches_doci_eigh(L, npair, g):
    """The eigenvalue-based Richardson-Gaudin solver against a dense eigh of the
    oracle's DOCI matrix of the RedBCS model (eps_i = i): the check that pins it
    before it serves as the exact energy of the L = 24 instance."""
    from oracle.richardson import redbcs_integrals, richardson_gaudin_energy
    eps = np.arange(L, dtype=float)
    E, _ = richardson_gaudin_energy(eps, npair, g)
    h, V = redbcs_integrals(eps, g)
    assert np.array_equal((h, V)[1], redbcs_ham(L, eps, g)[1])
    w = np.linalg.eigvalsh(O.dense_from_upper(O.build_crs("DOCI", h, V, npair, npair)))
    assert abs(E - w[0]) <= 1e-12 * max(1.0, abs(w[0]))


def test_richardson_gaudin_closed_forms():
    """One pair: E is the lowest root of 1 + g sum_i 1/(2 eps_i - E) = 0
    (Richardson's equation for N = 1); g -> 0: E -> sum of the N lowest 2 eps_i
    plus g N to first order; every level filled: E = 2 sum eps + g N (only the
    diagonal b_i^+ b_i of the pair hopping acts)."""
    from scipy.optimize import brentq

    from oracle.richardson import richardson_gaudin_energy
    eps = np.array([0.0, 0.7, 1.9, 2.4, 4.0])
    g = -0.35
    E, _ = richardson_gaudin_energy(eps, 1, g)
    f = lambda x: 1.0 + g * np.sum(1.0 / (2 * eps - x))
    E1 = brentq(f, -50.0, 2 * eps[0] - 1e-9)
    assert abs(E - E1) <= 1e-12
    gsmall = 1e-7
    E, _ = richardson_gaudin_energy(eps, 2, gsmall, steps=10)
    assert abs(E - (2 * (eps[0] + eps[1]) + gsmall * 2)) <= 1e-12
    E, _ = richardson_gaudin_energy(eps, 5, g, steps=50)
    assert abs(E - (2 * eps.sum() + g * 5)) <= 1e-11
