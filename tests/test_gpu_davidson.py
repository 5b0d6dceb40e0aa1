"""GPU tests of the Davidson solver (cf_solve_opts.method = CF_DAVIDSON), the
served solves' alternative to the ARPACK-shaped Lanczos.

* FP64 energies equal the Lanczos path's, the reference KATs and the
  oracle's within 1e-10 Eh, on stored CRS (FCI, DOCI, the block-staged DOCI
  kernel) and on the on-the-fly sigma; the returned vector is normalised and
  its residual is the one reported;
* FP32 values + FP64 polish with the Kato-Temple energy stop lands within
  1e-8 Eh of the FP64 energy;
* served with method=CF_DAVIDSON, every result is the same alone and in the
  mix, and within 1e-8 Eh of the oracle's FP64 solve.
"""
import json
import os
import sys

import numpy as np
import pytest

import ciflow_amd as CA
from ciflow_amd import stream as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def fixture(name):
    d = np.load(os.path.join(GOLD, name))
    return d["h"], d["V"], int(d["nalpha"]), int(d["nbeta"]), float(d["econst"])


def residual(cim, vec):
    y = cim.mvprod(vec) if isinstance(cim, CA.FCI_Direct) else cim.get_mat().mvprod(vec)
    th = float(vec @ y)
    return th, float(np.linalg.norm(y - th * vec))


@pytest.mark.parametrize("case", ["fci_L8_4a4b", "fci_L8_4a3b", "fci_L10_seed8", "doci_L12", "doci_L20_blk",
                                  "direct_L9", "fci_L11_large_n"])
def test_davidson_fp64_equals_lanczos(case):
    say("davidson", case)
    if case == "fci_L8_4a4b":
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(8, 0), 4, 4), CA.FCI
    elif case == "fci_L8_4a3b":
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(8, 3), 4, 3), CA.FCI
    elif case == "fci_L10_seed8":  # ground state outside the lowest determinant's spin sector
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(10, 8), 5, 5), CA.FCI
    elif case == "doci_L12":
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(12, 5), 6, 6), CA.DOCI
    elif case == "fci_L11_large_n":  # 213,444 determinants: the two-pass large-n step
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(11, 2), 5, 5), CA.FCI
    elif case == "doci_L20_blk":
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(20, 6), 10, 10), CA.DOCI
    else:
        ham, cls = CA.Hamiltonian(*CA.gen_integrals(9, 4), 4, 4), CA.FCI_Direct
    cim = cls(ham)
    ref = cim.solve()
    e_ref = cim.get_ci_energy()
    info = cim.solve(method=CA.CF_DAVIDSON)
    assert info.status == 0
    assert abs(cim.get_ci_energy() - e_ref) < 1e-10, (cim.get_ci_energy(), e_ref)
    vec = cim.get_eigvec()
    assert abs(np.linalg.norm(vec) - 1.0) < 1e-12
    th, res = residual(cim, vec)
    assert abs(th - info.eigval) < 1e-10
    assert res < 1e-10 * max(1.0, abs(th)) and abs(res - info.resid) < 1e-9
    say(case, "lanczos", ref.n_matvec, "davidson", info.n_matvec)


def test_davidson_reference_kat():
    """DOCI BeH2 STO-3G: -15.225362047972723 (TestDOCI.cpp:31)."""
    h, V, na, nb, ec = fixture("beh2_sto3g.npz")
    cim = CA.DOCI(CA.Hamiltonian(h, V, na, nb, ec))
    cim.solve(method=CA.CF_DAVIDSON)
    assert abs(cim.get_ci_energy() - (-15.225362047972723)) < 1e-10


@pytest.mark.parametrize("kind", ["FCI", "DOCI", "direct", "FCI_large_n"])
def test_davidson_fp32_polish_within_1e8(kind):
    say("davidson f32", kind)
    if kind == "FCI_large_n":
        ham = CA.Hamiltonian(*CA.gen_integrals(11, 5), 5, 5)
        cim = CA.FCI(ham, value_prec=CA.CF_F32)
    elif kind == "DOCI":
        ham = CA.Hamiltonian(*CA.gen_integrals(16, 8), 8, 8)
        cim = CA.DOCI(ham, value_prec=CA.CF_F32)
    elif kind == "FCI":
        ham = CA.Hamiltonian(*CA.gen_integrals(10, 8), 5, 5)
        cim = CA.FCI(ham, value_prec=CA.CF_F32)
    else:
        ham = CA.Hamiltonian(*CA.gen_integrals(10, 8), 5, 5)
        cim = CA.FCI_Direct(ham, value_prec=CA.CF_F32)
    cim.solve()
    e64 = cim.get_ci_energy()
    info = cim.solve(precision=CA.CF_F32, method=CA.CF_DAVIDSON, tol_energy=1e-9, check_every=2)
    assert info.n_matvec_f32 > 0 and info.n_matvec > 0
    assert abs(cim.get_ci_energy() - e64) < 1e-8


@pytest.mark.parametrize("method", [CA.CF_DAVIDSON, CA.CF_SOLVER_AUTO])
def test_served_davidson_alone_equals_in_mix_and_matches_oracle(method):
    reqs = S.make_stream(12, rate=0.0, doci_L=(12, 14), fci_L=(8, 9))
    arrs = S.prepare(reqs)
    kw = dict(precision=CA.CF_F32, check_every=2, tol_energy=1e-9, method=method, timeout=240)
    one = CA.SolveServer(slots=1)
    alone = []
    for r, a in zip(reqs, arrs):
        alone.append(S.run_stream([r], [a], server=one, **kw)["results"][0])
    one.close()
    mix = S.run_stream(reqs, arrs, slots=4, direct_fci=False, **kw)["results"]
    for a, b, r, arr in zip(alone, mix, reqs, arrs):
        assert a["status"] == 0 and b["status"] == 0
        assert a["energy"] == b["energy"] and a["n_matvec"] == b["n_matvec"], (r.kind, r.L)
        h, V, U = arr
        hr, Vr = O.rotate_integrals(U, h, V)
        m = O.build_crs("DOCI" if r.kind == CA.CF_DOCI else "FCI", hr, Vr, r.n, r.n)
        E, _, info = O.lanczos(O.full_from_upper(m), tol=1e-12)
        assert abs(b["energy"] - E) < 1e-8, (r.kind, r.L, b["energy"], E)


@pytest.mark.parametrize("method", [CA.CF_LANCZOS, CA.CF_DAVIDSON])
def test_resident_eigvec_read_back_with_sign_convention(method):
    """A solve without a host buffer leaves the vector on the device; reading
    it gives the same vector, sign convention included (largest-magnitude
    component positive), as a solve that returns it."""
    ham = CA.Hamiltonian(*CA.gen_integrals(9, 2), 4, 4)
    d = CA.FCI_Direct(ham)
    d.solve(want_vec=False, method=method)
    c = np.zeros(d.get_dim())
    CA._check(CA.lib().cf_space_eigvec(d.ptr, c.ctypes.data_as(CA.dp)))
    d.solve(want_vec=True, method=method)
    v = d.get_eigvec()
    assert np.array_equal(c, v)
    assert v[np.argmax(np.abs(v))] > 0


def test_davidson_fp32_phase_stops_at_its_noise_floor():
    """An FP32-phase tolerance the FP32 operator cannot reach (1e-14) ends the
    phase when its residual stalls, not at the iteration cap, and the FP64
    polish still lands on the FP64 energy."""
    ham = CA.Hamiltonian(*CA.gen_integrals(10, 3), 5, 5)
    cim = CA.FCI_Direct(ham, value_prec=CA.CF_F32)
    cim.solve()
    e64 = cim.get_ci_energy()
    info = cim.solve(precision=CA.CF_F32, method=CA.CF_DAVIDSON, tol_f32=1e-14)
    say("stall", info.n_matvec_f32, info.n_matvec)
    assert info.status == 0
    assert info.n_matvec_f32 < 2000
    assert abs(cim.get_ci_energy() - e64) < 1e-10


def test_solve_auto_picks_by_size():
    """method=CF_SOLVER_AUTO on a single solve: Davidson for a space of at
    least 2,000 determinants (FCI L=8 4a4b, 4,900), Lanczos below (DOCI L=10,
    252) -- the same products and energy as asking for that solver."""
    for cls, (L, n, seed), want in ((CA.FCI, (8, 4, 0), CA.CF_DAVIDSON), (CA.DOCI, (10, 5, 2), CA.CF_LANCZOS)):
        ham = CA.Hamiltonian(*CA.gen_integrals(L, seed), n, n)
        cim = cls(ham)
        auto = cim.solve(method=CA.CF_SOLVER_AUTO)
        e_auto = cim.get_ci_energy()
        ref = cim.solve(method=want)
        say("auto", cls.__name__, auto.n_matvec, ref.n_matvec)
        assert auto.n_matvec == ref.n_matvec and e_auto == cim.get_ci_energy()
