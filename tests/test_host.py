"""CPU tests of the GPU library's host side (no GPU needed):
  * the C-ABI library loads and exports every entry point include/ciflow_amd.h declares,
  * the synthetic integral generator is bit-identical to the oracle's,
  * the host reverse-communication Lanczos (cf_rci_*, the dsaupd_-shaped
    loop of SparseMatrix_CRS.cpp:441-449) finds the lowest eigenpair of
    numpy operators, including tiny spaces (n < ncv) and clustered spectra,
  * argument errors come back as status codes, not crashes.
"""
import os

import numpy as np
import pytest

import ciflow_amd as CA
from oracle import oracle as O


def test_library_exports_every_header_symbol():
    lib = CA.lib()
    syms = CA.header_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_version():
    assert CA.lib().cf_version() >= 100


@pytest.mark.parametrize("L,seed", [(4, 5), (8, 0), (12, 1)])
def test_generator_bit_identical_to_oracle(L, seed):
    h, V = CA.gen_integrals(L, seed)
    h2, V2 = O.gen_integrals(L, seed)
    assert np.array_equal(h, h2) and np.array_equal(V, V2)


def _rand_sym(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    A = (A + A.T) / 2
    return A * spread


@pytest.mark.parametrize("n,ncv", [(1, 42), (2, 42), (5, 42), (40, 42), (43, 42), (300, 42), (300, 8)])
def test_rci_lanczos_lowest_eigenpair(n, ncv):
    A = _rand_sym(n, n)
    w, U = np.linalg.eigh(A)
    ev, vec, info = CA.arpack_like(lambda x: A @ x, n, ncv=ncv, tol=1e-12)
    assert info == 0
    assert abs(ev - w[0]) < 1e-10 * max(1.0, abs(w[0]))
    assert abs(abs(vec @ U[:, 0]) - 1.0) < 1e-8
    # sign convention: largest-magnitude component positive
    assert vec[np.argmax(np.abs(vec))] > 0


def test_rci_lanczos_clustered_spectrum():
    n = 400
    rng = np.random.default_rng(3)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    d = np.concatenate([[-1.0, -1.0 + 1e-4], np.linspace(0, 5, n - 2)])
    A = (Q * d) @ Q.T
    ev, vec, info = CA.arpack_like(lambda x: A @ x, n, ncv=42, tol=1e-12)
    assert info == 0 and abs(ev - (-1.0)) < 1e-10


def test_rci_restart_limit_reports_info_1():
    A = _rand_sym(500, 9)
    r = CA.RCILanczos(500, ncv=6, tol=1e-14, maxrestart=1)
    while r.step() != 99:
        i0, i1 = r.ipntr[0] - 1, r.ipntr[1] - 1
        r.workd[i1:i1 + 500] = A @ r.workd[i0:i0 + 500]
    ev, vec, info = r.result()
    assert info == 1  # ARPACK's "maximum number of iterations reached"


def test_arguments_rejected_with_status_codes():
    lib = CA.lib()
    p = CA.vp()
    assert lib.cf_rci_create(10, 1, 1e-12, 0, 1, CA.C.byref(p)) == CA.CF_ERR_ARG
    assert b"ncv" in lib.cf_last_error()
    assert lib.cf_rci_create(0, 10, 1e-12, 0, 1, CA.C.byref(p)) == CA.CF_ERR_ARG
    assert lib.cf_gen_integrals(0, 1, None, None) == CA.CF_ERR_ARG
    assert lib.cf_ctx_create(0, None) != CA.CF_OK


# --- orbital rotation helpers (host side) ---------------------------------
def test_expm_skew_matches_scipy_and_is_orthogonal():
    import scipy.linalg as sl
    for L, seed in [(2, 1), (8, 2), (20, 3)]:
        K = CA.gen_rotation(L, seed, 0.05)
        assert np.array_equal(K, -K.T) and np.all(np.diag(K) == 0)
        U = CA.expm_skew(K)
        assert np.max(np.abs(U - sl.expm(K))) < 1e-14
        assert np.max(np.abs(U @ U.T - np.eye(L))) < 1e-14
        assert abs(np.linalg.det(U) - 1.0) < 1e-13
        assert np.max(np.abs(CA.expm_skew(-K) @ U - np.eye(L))) < 1e-14
    # 2x2 closed form: exp([[0, t], [-t, 0]]) = [[cos t, sin t], [-sin t, cos t]]
    t = 0.7
    U = CA.expm_skew(np.array([[0.0, t], [-t, 0.0]]))
    assert np.max(np.abs(U - np.array([[np.cos(t), np.sin(t)], [-np.sin(t), np.cos(t)]]))) < 1e-15
    # a large rotation exercises the squaring phase
    K = CA.gen_rotation(6, 9, 1.5)
    assert np.max(np.abs(CA.expm_skew(K) - sl.expm(K))) < 1e-12


# --- served stream: host side ------------------------------------------------
def test_stream_spec_is_deterministic_and_mixed():
    from ciflow_amd import stream as S
    a = S.make_stream(2000, rate=100.0)
    b = S.make_stream(2000, rate=100.0)
    assert a == b
    doci = [r for r in a if r.kind == CA.CF_DOCI]
    fci = [r for r in a if r.kind == CA.CF_FCI]
    assert abs(len(doci) / len(a) - 0.7) < 0.04
    assert {r.L for r in doci} == set(range(12, 21)) and {r.L for r in fci} == set(range(8, 12))
    assert all(r.n == r.L // 2 and r.seed == 1000 + r.k for r in a)
    t = np.array([r.t_arrive for r in a])
    assert t[0] == 0.0 and np.all(np.diff(t) >= 0)
    assert abs(t[-1] / (len(a) - 1) - 0.01) < 0.001   # mean gap 1/rate
    h, V, U = S.request_arrays(a[0])
    assert np.max(np.abs(U @ U.T - np.eye(a[0].L))) < 1e-14


def test_request_footprint_and_argument_checks():
    fp = CA.SolveServer.footprint
    small = fp(CA.CF_DOCI, 12, 6, 6, CA.CF_F32)
    big = fp(CA.CF_FCI, 11, 5, 5, CA.CF_F32)
    assert 64 << 20 < small < big
    # FCI L=11 5a5b: 213,444 rows of 1,261 entries at 16 B (f64 + f32 value, u32 column)
    assert 4.2e9 < big < 4.6e9
    assert fp(CA.CF_FCI, 11, 5, 5, CA.CF_F64) < big
    # FCI L=10 5a5b (252 alpha strings, no position-major copy): 63,504 rows of
    # 876 entries, stored twice (CRS and the sliced copy)
    assert fp(CA.CF_FCI, 10, 5, 5, CA.CF_F64) > 2 * 63504 * 876 * 12
    with pytest.raises(CA.CIFlowError):
        fp(CA.CF_FCI, 8, 9, 4)


def test_determinant_file_parser_matches_oracle(tmp_path):
    """Permutator_File::read_file semantics: '#' lines are comments, strings are
    binary with the highest orbital first, CR line ends are tolerated, lines
    without '|' are skipped, strings beyond 64 orbitals are rejected."""
    gold = os.path.join(os.path.dirname(__file__), "golden")
    for name in ("GKCI_Be_cntSp_cc-pVDZ_0100.txt", "fci_beh2_sto-3gdeterminants.dat"):
        path = os.path.join(gold, name)
        up, down = CA.read_determinants(path)
        ou, od = O.read_dets(path)
        assert np.array_equal(up, ou) and np.array_equal(down, od)
    p = tmp_path / "dets.txt"
    p.write_bytes(b"# header | with a bar\n0011|0101\r\n\n1001|0011\n#0110|0110\n")
    up, down = CA.read_determinants(str(p))
    assert up.tolist() == [3, 9] and down.tolist() == [5, 3]
    p.write_text("1" * 65 + "|" + "1" * 65 + "\n")
    with pytest.raises(CA.CIFlowError):
        CA.read_determinants(str(p))


def test_integral_file_reader_matches_fixture(tmp_path):
    """Hamiltonian::read_file on the reference's BeH2 dump gives, bit for bit,
    the arrays of the committed fixture (converted by an independent reader)."""
    gold = os.path.join(os.path.dirname(__file__), "golden")
    d = CA.read_integrals(os.path.join(gold, "beh2_sto3g_1_5.dat"))
    f = np.load(os.path.join(gold, "beh2_sto3g.npz"))
    assert (d["L"], d["nalpha"], d["nbeta"]) == (int(f["L"]), int(f["nalpha"]), int(f["nbeta"]))
    assert d["econst"] == float(f["econst"])
    assert np.array_equal(d["h"], f["h"]) and np.array_equal(d["V"], f["V"])
    assert np.array_equal(d["irreps"], f["irreps"])
    with pytest.raises(CA.CIFlowError):
        CA.read_integrals(str(tmp_path / "missing.dat"))
    p = tmp_path / "cut.dat"
    p.write_text("header\n****  Molecular Integrals For DOCI Start Here\nNalpha = 1\n")
    with pytest.raises(CA.CIFlowError):
        CA.read_integrals(str(p))
