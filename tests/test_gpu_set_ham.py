"""CIMethod::set_ham with resident matrices (cf_space_set_ham): new integrals
for the same space refill the values in place and must give exactly what a
fresh build of the same integrals gives -- CRS rows bit-exact, products within
rounding, energies within 1e-10 Eh -- including when an element crosses the
drop threshold and the pattern changes."""
import numpy as np
import pytest

import ciflow_amd as CA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def ham_pair(L, na, nb, seed):
    """Two Hamiltonians of one shape: seeded integrals, and other seeded
    integrals rotated by exp(K) (an FCI energy is invariant under the
    rotation alone)."""
    h, V = CA.gen_integrals(L, seed)
    a = CA.Hamiltonian(h, V, na, nb)
    h2, V2 = CA.gen_integrals(L, seed + 1)
    U = CA.expm_skew(CA.gen_rotation(L, seed + 100))
    return a, CA.Hamiltonian(h2, V2, na, nb).rotate(U)


def same_rows(m1, m2):
    r1, c1, v1 = m1.export_full()
    r2, c2, v2 = m2.export_full()
    assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
    assert np.array_equal(v1, v2)


@pytest.mark.parametrize("kind,L,na,nb", [("DOCI", 10, 5, 5), ("DOCI", 18, 9, 9), ("DOCI", 20, 10, 10),
                                          ("FCI", 9, 4, 4),
                                          ("FCI", 11, 5, 5), ("FCI", 8, 4, 3)])
def test_set_ham_equals_fresh_build(kind, L, na, nb):
    """Resident matrix after set_ham == a matrix built for the new integrals:
    position-major copies (FCI L=11 kernel 4, DOCI L=18 kernel 5, DOCI L=20
    kernel 6) refilled,
    CRS rebuilt on first read, FP64 and FP32 values."""
    say("set_ham", kind, L, na, nb)
    cls = CA.FCI if kind == "FCI" else CA.DOCI
    a, b = ham_pair(L, na, nb, 7)
    cim = cls(a, value_prec=CA.CF_F32)
    cim.solve()
    ea = cim.get_ci_energy()
    cim.set_ham(b)
    assert cim.get_eigvec() is None
    fresh = cls(b, value_prec=CA.CF_F32)
    m, mf = cim.get_mat(), fresh.get_mat()
    x = np.random.default_rng(3).standard_normal(cim.get_dim())
    for prec in (CA.CF_F64, CA.CF_F32):
        assert m.spmv_bytes(prec)[0] == mf.spmv_bytes(prec)[0]
        y, yf = m.mvprod(x, prec), mf.mvprod(x, prec)
        assert np.max(np.abs(y - yf)) <= 1e-13 * np.max(np.abs(yf)), prec
    cim.solve()
    fresh.solve()
    assert abs(cim.get_ci_energy() - fresh.get_ci_energy()) < 1e-10
    assert abs(cim.get_ci_energy() - ea) > 1e-6  # the integrals really changed
    same_rows(m, mf)
    assert m.info() == mf.info()
    # and back: the original integrals give the original energy
    cim.set_ham(a)
    cim.solve()
    assert abs(cim.get_ci_energy() - ea) < 1e-10


def test_set_ham_pattern_change_rebuilds_crs():
    """An element that drops below the threshold changes the CRS pattern:
    the rebuilt rows and nonzero count equal the fresh build's; the pair
    table of the DOCI hop kernel zeroes the same element."""
    L, n = 18, 9
    h, V = CA.gen_integrals(L, 4)
    a = CA.Hamiltonian(h, V, n, n)
    V2 = V.copy()
    for (j, k) in ((0, 5), (3, 11)):  # pair elements V(m,m,M,M) of two hops
        V2[j, j, k, k] = V2[k, k, j, j] = 0.0
    b = CA.Hamiltonian(h, V2, n, n)
    cim = CA.DOCI(a)
    nnz_a = cim.get_mat().info()[1]
    cim.set_ham(b)
    fresh = CA.DOCI(b)
    assert cim.get_mat().info() == fresh.get_mat().info()
    assert cim.get_mat().info()[1] < nnz_a
    same_rows(cim.get_mat(), fresh.get_mat())
    x = np.random.default_rng(1).standard_normal(cim.get_dim())
    for kern in (5, 6, 1, 3):
        cim.get_mat().set_spmv_kernel(kern)
        y = cim.get_mat().mvprod(x)
        yf = fresh.get_mat().mvprod(x)
        assert np.max(np.abs(y - yf)) <= 1e-13 * np.max(np.abs(yf)), kern


def test_set_ham_direct_and_det_list_and_new_shape():
    """FCI_Direct remakes its sigma tables; FCI_File keeps its list; another
    electron count gets a new space."""
    a, b = ham_pair(8, 4, 4, 11)
    d = CA.FCI_Direct(a)
    d.set_ham(b)
    d.solve()
    ref = CA.FCI(b)
    ref.solve()
    assert abs(d.get_ci_energy() - ref.get_ci_energy()) < 1e-10
    # a sigma made before the swap refuses to run on the new integrals
    old = CA.FCI_Direct(a)
    CA._check(CA.lib().cf_space_set_ham(old.ptr, b.ptr))
    old.ham = b
    old.space.parent = b.handle
    with pytest.raises(CA.CIFlowError):
        old.mvprod(np.ones(old.get_dim()))
    # selected CI over a fixed list
    strs = CA.default_context().unrank(8, 4, 0, 40)
    dets = (np.repeat(strs, 40), np.tile(strs, 40))
    sel = CA.FCI_File(a, dets)
    sel.set_ham(b)
    sel.solve()
    sel_ref = CA.FCI_File(b, dets)
    sel_ref.solve()
    assert abs(sel.get_ci_energy() - sel_ref.get_ci_energy()) < 1e-10
    same_rows(sel.get_mat(), sel_ref.get_mat())
    # another shape: a new space
    h, V = CA.gen_integrals(8, 11)
    c = CA.Hamiltonian(h, V, 3, 3)
    f = CA.FCI(a)
    f.set_ham(c)
    assert f.get_dim() == 56 * 56
    f.solve()
    g = CA.FCI(c)
    g.solve()
    assert abs(f.get_ci_energy() - g.get_ci_energy()) < 1e-10
