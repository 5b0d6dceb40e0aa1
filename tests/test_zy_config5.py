"""GPU test of the on-the-fly sigma at full size: FCI L=16 8a8b, seeded
integrals (seed 3), 165,636,900 determinants -- the role of the reference's
CI_Big (CIMethod.h:176-189, CIMethod.cpp:943-1116), whose O(n^2) mvprod
cannot reach this size.

The FP32-sigma Davidson phase plus its FP64 refinement runs to a residual
tolerance of 5e-9 |theta|, then:
* the returned vector is an eigenvector of the FP64 operator: the residual
  |H x - theta x| formed with one more FP64 sigma is <= 1e-8 |E|;
* the energy equals the FP64 Lanczos energy of the same operator,
  -44.926334965054 Eh (a 300-product solve), within 1e-8 Eh;
* the 1-RDM of the vector has trace 16 (CIDens.cpp:365-409 invariant).
The sigma's rows themselves are pinned against the oracle's Slater-Condon
elements in test_gpu_parity.py.  Collected next to last (a two-minute solve);
prints its progress.
"""
import sys
import time

import numpy as np
import pytest

import ciflow_amd as CA

pytestmark = pytest.mark.gpu

E_FP64_LANCZOS = -44.926334965054


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_config5_fp32_refinement_residual_and_rdm_trace():
    L, nel, seed = 16, 8, 3
    h, V = CA.gen_integrals(L, seed)
    t0 = time.perf_counter()
    d = CA.FCI_Direct(CA.Hamiltonian(h, V, nel, nel), value_prec=CA.CF_F32)
    say(f"config 5: n = {d.n}, plan {time.perf_counter() - t0:.1f} s")
    assert d.n == 165636900
    last = [0.0]

    def prog(nmv, th, r):
        if time.perf_counter() - last[0] > 10.0:
            last[0] = time.perf_counter()
            say(f"  {nmv} products, theta {th:.12f}, residual {r:.2e}, {time.perf_counter() - t0:.0f} s")

    info = d.solve(precision=CA.CF_F32, method=CA.CF_DAVIDSON, ncv=39, tol=5e-9, progress=prog)
    E = d.get_ci_energy()
    say(f"  E = {E:.12f}, {info.n_matvec_f32} FP32 + {info.n_matvec} FP64 products, "
        f"{time.perf_counter() - t0:.0f} s")
    assert info.status == 0
    assert abs(E - E_FP64_LANCZOS) <= 1e-8, E
    x = np.asarray(d.get_eigvec(), dtype=np.float64)
    assert abs(np.linalg.norm(x) - 1.0) <= 1e-10
    y = d.mvprod(x, precision=CA.CF_F64)
    res = float(np.linalg.norm(y - info.eigval * x))
    del y
    say(f"  |H x - theta x| = {res:.3e} (bound {1e-8 * abs(E):.3e})")
    assert res <= 1e-8 * abs(E)
    d.construct_density(trdm=False)
    tr = sum(d.get_one_rdm(s, i, i) for s in range(2) for i in range(L))
    say(f"  tr gamma = {tr:.12f}")
    assert abs(tr - 2 * nel) <= 1e-9
