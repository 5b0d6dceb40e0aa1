"""Exact ground-state energy of the reduced-BCS (Richardson) pairing model --
test infrastructure: the oracle for the RedBCS DOCI instance, imported only by
tests/.

The model is the reference's RedBCS (RedBCS.cpp:38-121, ModHam.cpp:174-205):
h_ii = eps_i, (ii|jj) = g for all i, j, _modham set, solved as DOCI.  On the
seniority-zero space of N pairs in L levels the DOCI matrix is
    H_II = sum_{i occupied} 2 eps_i + g N,     H_IJ = g (one pair hop),
i.e. H = sum_i e_i b_i^+ b_i + g sum_{ij} b_i^+ b_j with e_i = 2 eps_i.

Richardson (R. W. Richardson, Phys. Lett. 3, 277 (1963)): E = sum_a E_a with
the N pair energies solving
    1/g + sum_i 1/(e_i - E_a) + 2 sum_{b != a} 1/(E_a - E_b) = 0.        (R)
(one pair: 1 + g sum_i 1/(e_i - E) = 0, the secular equation of H above).
These equations are singular where pair energies collide (at g = -0.5 and
e_i = 2i several pairs have turned complex), so this solver uses the
eigenvalue-based variables of Faribault, El Araby, Straeter and Gritsev
(Phys. Rev. B 83, 235124 (2011)):  Lambda_j = sum_a 1/(e_j - E_a), real and
smooth in g along the ground state.  With Lambda(z) = sum_a 1/(z - E_a),
Lambda^2 + Lambda' = 2 sum_a [sum_{b!=a} 1/(E_a - E_b)] / (z - E_a); inserting
(R) and letting z -> e_j gives, for every level j,
    Lambda_j^2 + Lambda_j / g + sum_{i != j} (Lambda_j - Lambda_i)/(e_i - e_j) = 0,
and summing e_j Lambda_j with (R) gives the energy
    E = g (N (L - N + 1) - sum_j e_j Lambda_j).
In u_j = g Lambda_j the equations read
    u_j^2 + u_j + g sum_{i != j} (u_j - u_i)/(e_i - e_j) = 0,   E = g N (L - N + 1) - sum_j e_j u_j,
and at g = 0 the ground state is u_j = -1 on the N lowest levels and 0 above
(E = sum of the N lowest e_j).  The solver follows that root from g = 0 to the
target g by Newton steps on a fine grid of g (continuation), each started from
the quadratic extrapolation of the last three solutions.
Levels must be distinct.
"""
from __future__ import annotations

import numpy as np


def _residual(u, e, g):
    d = e[None, :] - e[:, None]          # d[j, i] = e_i - e_j
    np.fill_diagonal(d, np.inf)
    w = 1.0 / d                          # 0 on the diagonal
    F = u * u + u + g * ((u[:, None] - u[None, :]) * w).sum(axis=1)
    J = -g * w
    J[np.diag_indices_from(J)] = 2.0 * u + 1.0 + g * w.sum(axis=1)
    return F, J


def richardson_gaudin_energy(eps, npair: int, g: float, steps: int = 2000, tol: float = 1e-14):
    """Ground-state energy of the reduced-BCS model h_ii = eps_i, (ii|jj) = g
    with `npair` pairs; returns (E, u) with u_j = g Lambda_j."""
    eps = np.asarray(eps, dtype=np.float64)
    L = len(eps)
    if not 0 <= npair <= L:
        raise ValueError("0 <= npair <= L")
    e = 2.0 * eps
    if len(np.unique(e)) != L:
        raise ValueError("levels must be distinct")
    order = np.argsort(e)
    u = np.zeros(L)
    u[order[:npair]] = -1.0
    hist = [(0.0, u.copy())]
    gs = np.linspace(0.0, g, steps + 1)[1:]
    for gk in gs:
        if len(hist) >= 3:   # quadratic extrapolation in g
            (g0, u0), (g1, u1), (g2, u2) = hist[-3:]
            l0 = (gk - g1) * (gk - g2) / ((g0 - g1) * (g0 - g2))
            l1 = (gk - g0) * (gk - g2) / ((g1 - g0) * (g1 - g2))
            l2 = (gk - g0) * (gk - g1) / ((g2 - g0) * (g2 - g1))
            u = l0 * u0 + l1 * u1 + l2 * u2
        for _ in range(50):
            F, J = _residual(u, e, gk)
            du = np.linalg.solve(J, -F)
            u = u + du
            if np.max(np.abs(du)) <= tol * max(1.0, np.max(np.abs(u))):
                break
        else:
            raise RuntimeError(f"Newton did not converge at g = {gk}")
        hist.append((gk, u.copy()))
        hist = hist[-3:]
    F, _ = _residual(u, e, g)
    if np.max(np.abs(F)) > 1e-10:
        raise RuntimeError(f"residual {np.max(np.abs(F))} at the target g")
    E = g * npair * (L - npair + 1) - float(e @ u)
    return E, u


def redbcs_integrals(eps, g: float):
    """h, V of the RedBCS model as the reference builds them (RedBCS::construct_ham,
    ModHam.cpp:194-205) and as getVmat returns them with _modham set
    (Hamiltonian.cpp:311-318): h_ii = eps_i, V[i][j][k][l] = g when i == j and
    k == l (setVmat(i,i,j,j,g) for all i, j), zero otherwise."""
    eps = np.asarray(eps, dtype=np.float64)
    L = len(eps)
    h = np.diag(eps)
    V = np.zeros((L, L, L, L))
    for i in range(L):
        V[i, i, :, :][np.arange(L), np.arange(L)] = g
    return h, V
