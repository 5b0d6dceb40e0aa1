"""Orbital-optimisation callers of the GPU solver.

The reference's optimisers (LocalMinimizer.cpp:73-235, SimulatedAnnealing.cpp
:83-87, 274-390) repeat update_cim -> set_ham -> build_ci -> solve for many
rotated Hamiltonians.  Here every such step is one GPU rotation of the original
integrals by the accumulated orthogonal matrix (FP64 MFMA), one GPU build and
solve, and -- for the local minimiser -- the GPU RDMs, while the decisions stay
on the host as in the reference:

  * jacobi_rotation: the Givens rotation of OrbitalTransform::do_jacobi_rotation
    (OrbitalTransform.cpp:278-330; new k = cos k - sin l, new l = cos l + sin k).
  * DociRotationModel: the DOCI energy of the current wavefunction after a
    rotation of orbitals k, l, as a trigonometric polynomial in the angle whose
    coefficients are contractions of the 1-/2-RDM with the integrals
    (DensDOCI::calc_rotate / find_min_angle, CIDens.cpp:1279-1464), and its
    Newton search for the minimising angle.
  * LocalMinimizer: scan every orbital pair with the model, rotate the best
    pair, solve again; stop after conv_steps energy changes below conv_crit.
  * SimulatedAnnealing: Metropolis walk over random Jacobi rotations with a
    seeded generator (the reference draws from std::random_device).
  * OrbitalOptimization.energies_of: a batch of candidate rotations solved
    concurrently by the served stream (cf_server), the route for callers that
    evaluate many independent rotations.
"""
from __future__ import annotations

import math

import numpy as np

from . import CF_DOCI, CF_F64, CF_FCI, DOCI, FCI, Hamiltonian, SolveServer


def jacobi_rotation(L: int, k: int, l: int, theta: float) -> np.ndarray:
    """U (new orbitals in rows) of do_jacobi_rotation(k, l, theta)."""
    U = np.eye(L)
    c, s = math.cos(theta), math.sin(theta)
    U[k, k], U[k, l] = c, -s
    U[l, l], U[l, k] = c, s
    return U


class DociRotationModel:
    """E(theta) = base + A c^4 + B s^4 + C c^2 + D s^2 + 2 E s c + 2 F c^2 s^2
    + 4 G c^3 s + 4 H c s^3 for a rotation of orbitals k, l of a DOCI state,
    with c = cos(theta), s = sin(theta); the coefficients contract the
    integrals with the state's RDMs (CIDens.cpp:1279-1464).  G0/G1 are the
    alpha-alpha and alpha-beta 2-RDM blocks as dense [i][j][k][l] arrays, so
    G1[k,k,l,l] is the pair-hop element and G0[k,a,k,a] the same-spin pair
    density."""

    def __init__(self, h, V, G0, G1, nelec: int):
        self.T = np.asarray(h)
        self.V = np.asarray(V)
        self.G0 = np.asarray(G0)
        self.G1 = np.asarray(G1)
        self.f = 2.0 / (nelec - 1.0)
        L = self.T.shape[0]
        self.L = L
        idx = np.arange(L)
        self.Td = self.T[idx, idx]
        self.VJ = self.V[idx[:, None], idx[:, None], idx[None, :], idx[None, :]]  # V(a,a,b,b)
        self.VC = self.V[idx[:, None], idx[None, :], idx[:, None], idx[None, :]]  # V(a,b,a,b)
        self.VX = self.V[idx[:, None], idx[None, :], idx[None, :], idx[:, None]]  # V(a,b,b,a)
        self.P = self.G1[idx[:, None], idx[:, None], idx[None, :], idx[None, :]]  # G1(a,a,b,b)
        self.D = self.G0[idx[:, None], idx[None, :], idx[:, None], idx[None, :]]  # G0(a,b,a,b)

    def coefficients(self, k: int, l: int):
        T, V, P, D, f = self.T, self.V, self.P, self.D, self.f
        a = np.array([x for x in range(self.L) if x != k and x != l], dtype=int)
        Pkk, Pll, Pkl = P[k, k], P[l, l], P[k, l]
        Dkl = D[k, l]
        wk = 2 * self.VC[k, a] - self.VX[k, a] + f * T[k, k]
        wl = 2 * self.VC[l, a] - self.VX[l, a] + f * T[l, l]
        wkl = 2 * V[k, a, l, a] - V[k, a, a, l] + f * T[k, l]
        C = f * (T[k, k] * Pkk + T[l, l] * Pll) + np.sum(
            2 * self.VJ[k, a] * P[k, a] + 2 * self.VJ[l, a] * P[l, a] + 2 * wk * D[k, a] + 2 * wl * D[l, a])
        Dc = f * (T[l, l] * Pkk + T[k, k] * Pll) + np.sum(
            2 * self.VJ[l, a] * P[k, a] + 2 * self.VJ[k, a] * P[l, a] + 2 * wk * D[l, a] + 2 * wl * D[k, a])
        E = f * T[k, l] * (Pll - Pkk) + np.sum(
            2 * V[k, l, a, a] * (P[l, a] - P[k, a]) + 2 * wkl * (D[l, a] - D[k, a]))
        Vkkkk, Vllll = V[k, k, k, k], V[l, l, l, l]
        Vkkll, Vklkl = V[k, k, l, l], V[k, l, k, l]
        Vklkk, Vklll = V[k, l, k, k], V[k, l, l, l]
        A = Vkkkk * Pkk + Vllll * Pll + 2 * Vkkll * Pkl + 2 * (2 * Vklkl - Vkkll) * Dkl
        B = Vkkkk * Pll + Vllll * Pkk + 2 * Vkkll * Pkl + 2 * (2 * Vklkl - Vkkll) * Dkl
        F = ((2 * Vkkll + Vklkl) * (Pkk + Pll) + (Vkkkk + Vllll - 2 * (Vklkl + Vkkll)) * Pkl
             + (Vkkkk + Vllll - 6 * Vkkll + 2 * Vklkl) * Dkl)
        H = Vklkk * Pll - Vklll * Pkk - (Vklkk - Vklll) * (Pkl + Dkl)
        G = Vklll * Pll - Vklkk * Pkk + (Vklkk - Vklll) * (Pkl + Dkl)
        return A, B, C, Dc, E, F, G, H

    def base(self, k: int, l: int) -> float:
        """The angle-independent part (terms without k or l, and the k-l pair)."""
        T, f = self.T, self.f
        keep = np.array([x for x in range(self.L) if x != k and x != l], dtype=int)
        a = keep[:, None]
        b = keep[None, :]
        e = 2 * f * (T[k, k] + T[l, l]) * self.D[k, l]
        Ta = self.Td[keep]
        e += np.sum(f * Ta * (self.P[keep, keep] + 2 * self.D[keep, k] + 2 * self.D[keep, l]))
        e += np.sum(f * (self.Td[a] + self.Td[b]) * self.D[a, b] + self.VJ[a, b] * self.P[a, b]
                    + (2 * self.VC[a, b] - self.VX[a, b]) * self.D[a, b])
        return float(e)

    def energy(self, k: int, l: int, theta: float) -> float:
        """DensDOCI::calc_rotate (without Econst)."""
        A, B, C, D, E, F, G, H = self.coefficients(k, l)
        c, s = math.cos(theta), math.sin(theta)
        return float(self.base(k, l) + c ** 4 * A + s ** 4 * B + c * c * C + s * s * D + 2 * s * c * E
                     + 2 * c * c * s * s * F + 4 * c * s ** 3 * H + 4 * c ** 3 * s * G)

    @staticmethod
    def _grad(cf, theta):
        A, B, C, D, E, F, G, H = cf
        c, s = math.cos(theta), math.sin(theta)
        return (16 * (G - H) * c ** 4 - 4 * (A + B - 2 * F) * s * c ** 3 + 4 * (E - 3 * G + 5 * H) * c * c
                + 2 * (2 * B - C + D - 2 * F) * s * c - 2 * E - 4 * H)

    @staticmethod
    def _hess(cf, theta):
        A, B, C, D, E, F, G, H = cf
        c, s = math.cos(theta), math.sin(theta)
        return (-16 * (A + B - 2 * F) * c ** 4 + 64 * (H - G) * s * c ** 3 + 4 * (3 * A + 5 * B - C + D - 8 * F) * c * c
                - 8 * (E - 3 * G + 5 * H) * s * c + 2 * (C - 2 * B - D + 2 * F))

    def gradient(self, k: int, l: int, theta: float) -> float:
        return float(self._grad(self.coefficients(k, l), theta))

    def hessian(self, k: int, l: int, theta: float) -> float:
        return float(self._hess(self.coefficients(k, l), theta))

    def find_min_angle(self, k: int, l: int, start: float = 0.0, max_iters: int = 20, conv: float = 1e-12):
        """DensDOCI::find_min_angle: Newton on dE/dtheta from `start` (its sign
        flipped first when the quadratic model goes uphill); returns (angle,
        True for a minimum)."""
        cf = self.coefficients(k, l)
        theta = start
        change = self._grad(cf, theta) * theta + self._hess(cf, theta) * theta * theta / 2.0
        if change > 0:
            theta = -theta
        for _ in range(max_iters):
            dx = self._grad(cf, theta) / self._hess(cf, theta)
            theta -= dx
            if abs(dx) < conv:
                break
        return theta, self._hess(cf, theta) > 0


class OrbitalOptimization:
    """OrbitalOptimization (SimulatedAnnealing.cpp:83-87): the original
    integrals, the accumulated orthogonal matrix U (new orbitals in rows) and
    the CI method re-solved on the GPU for the rotated Hamiltonian."""

    def __init__(self, h, V, nup: int, ndown: int, econst: float = 0.0, kind: int = CF_DOCI,
                 precision: int = CF_F64, ctx=None):
        self.ham0 = Hamiltonian(h, V, nup, ndown, econst, ctx)
        self.kind = kind
        self.precision = precision
        self.L = self.ham0.L
        self.U = np.eye(self.L)
        self.cim = None
        self.ham = None
        self.update_cim()

    def update_cim(self) -> float:
        """set_ham + build_ci + solve for the current U; returns the CI energy.
        The CI method keeps its space and matrix resident across updates
        (CIMethod.set_ham refills the values in place)."""
        self.ham = self.ham0.rotate(self.U)
        if self.cim is None:
            cls = DOCI if self.kind == CF_DOCI else FCI
            self.cim = cls(self.ham, value_prec=self.precision)
        else:
            self.cim.set_ham(self.ham)
        self.cim.solve(precision=self.precision)
        return self.energy()

    def energy(self) -> float:
        return self.cim.get_ci_energy()

    def apply_rotation(self, k: int, l: int, theta: float) -> float:
        """do_jacobi_rotation on top of the current orbitals, then update_cim."""
        self.U = jacobi_rotation(self.L, k, l, theta) @ self.U
        return self.update_cim()

    def energies_of(self, rotations, slots: int = 8, precision: int | None = None, tol_energy: float = 0.0):
        """CI energies of the original Hamiltonian rotated by each U in
        `rotations` (new orbitals in rows), solved concurrently on the served
        stream; returns a list in input order."""
        prec = self.precision if precision is None else precision
        h, V = self.ham0.h, self.ham0.V
        out = [None] * len(rotations)
        with SolveServer(slots=slots) as srv:
            for t, U in enumerate(rotations):
                srv.submit(t, self.kind, h, V, self.ham0.nup, self.ham0.ndown, U=np.asarray(U),
                           econst=self.ham0.econst, precision=prec, tol_energy=tol_energy,
                           check_every=2 if tol_energy > 0 else 0)
            done = 0
            while done < len(rotations):
                for d in srv.poll():
                    if d["status"] != 0:
                        raise RuntimeError(f"rotation {d['tag']} failed: {d['error']}")
                    out[d["tag"]] = d["energy"]
                    done += 1
        return out


class LocalMinimizer(OrbitalOptimization):
    """LocalMinimizer (LocalMinimizer.cpp:22-235) for DOCI: each iteration
    models every orbital pair with the RDMs of the current solution, rotates
    the pair of lowest model energy (not the same pair twice in a row) and
    solves again.  conv_crit = 1e-6 and conv_steps = 25 as the reference."""

    def __init__(self, h, V, nup: int, ndown: int, econst: float = 0.0, precision: int = CF_F64, ctx=None,
                 conv_crit: float = 1e-6, conv_steps: int = 25, seed: int = 0):
        self.conv_crit = conv_crit
        self.conv_steps = conv_steps
        self.rng = np.random.default_rng(seed)
        super().__init__(h, V, nup, ndown, econst, CF_DOCI, precision, ctx)
        self.history = []

    def model(self) -> DociRotationModel:
        _, G = self.cim.density_dense()
        return DociRotationModel(self.ham.h, self.ham.V, G[0], G[1], self.ham.nup + self.ham.ndown)

    def scan_orbitals(self):
        """LocalMinimizer::scan_orbitals: [(k, l, angle, model energy + Econst)]."""
        m = self.model()
        out = []
        for k in range(self.L):
            for l in range(k + 1, self.L):
                th, ok = m.find_min_angle(k, l, 0.3)
                if not ok:
                    th, ok = m.find_min_angle(k, l, 0.01)
                if not ok or abs(th) > math.pi / 2:
                    continue
                out.append((k, l, th, m.energy(k, l, th) + self.ham.econst))
        return out

    def choose_orbitalpair(self, rots, e_ref):
        """LocalMinimizer::choose_orbitalpair: pick a pair with probability
        proportional to its energy gain."""
        w = np.array([e_ref - r[3] for r in rots])
        cum = np.cumsum(w / w.sum())
        return int(min(np.searchsorted(cum, self.rng.random(), side="right"), len(rots) - 1))

    def optimize(self, dist_choice: bool = False, max_iters: int = 1000) -> float:
        converged = 0
        opt_energy = 0.0
        prev = (0, 0)
        iters = 1
        while converged < self.conv_steps:
            rots = sorted(self.scan_orbitals(), key=lambda r: r[3])
            if not rots:
                break
            idx = self.choose_orbitalpair(rots, opt_energy) if dist_choice else 0
            if (rots[idx][0], rots[idx][1]) == prev and idx + 1 < len(rots):
                idx += 1
            k, l, th, e_model = rots[idx]
            prev = (k, l)
            e = self.apply_rotation(k, l, th)
            self.history.append((k, l, th, e_model, e))
            if abs(opt_energy - e) < self.conv_crit:
                converged += 1
            opt_energy = e
            iters += 1
            if iters > max_iters:
                break
        return self.energy()


class SimulatedAnnealing(OrbitalOptimization):
    """SimulatedAnnealing::optimize (SimulatedAnnealing.cpp:306-390): random
    Jacobi rotations with angle max_angle * (u1 - u2), Metropolis acceptance
    (SimulatedAnnealing.cpp:274-295), temperature and maximal angle shrunk by
    the given factors each step; the generator is seeded."""

    def __init__(self, h, V, nup: int, ndown: int, econst: float = 0.0, kind: int = CF_DOCI,
                 precision: int = CF_F64, ctx=None, temp: float = 1e-3, change_temp: float = 0.99,
                 max_angle: float = 0.5, change_angle: float = 0.99, steps: int = 200, seed: int = 0):
        self.temp, self.change_temp = temp, change_temp
        self.max_angle, self.change_angle = max_angle, change_angle
        self.steps = steps
        self.rng = np.random.default_rng(seed)
        super().__init__(h, V, nup, ndown, econst, kind, precision, ctx)
        self.opt_energy = self.energy()
        self.opt_U = self.U.copy()

    def accept(self, e_new: float, e_old: float, temp: float) -> bool:
        if e_new < e_old:
            return True
        cost = math.exp((e_old - e_new) / temp)
        return not (self.rng.random() * (1 + cost) > cost)

    def optimize(self) -> float:
        temp, max_angle = self.temp, self.max_angle
        e_old = self.energy()
        unaccept = 0
        for _ in range(self.steps):
            i = int(self.rng.integers(0, self.L))
            j = int(self.rng.integers(0, self.L))
            if i == j:
                continue
            angle = max_angle * (self.rng.random() - self.rng.random())
            U_old = self.U
            e_new = self.apply_rotation(i, j, angle)
            if self.accept(e_new, e_old, temp):
                e_old = e_new
                if e_new < self.opt_energy:
                    self.opt_energy = e_new
                    self.opt_U = self.U.copy()
                unaccept = 0
            else:
                self.U = U_old  # back to the previous orbitals
                self.update_cim()
                unaccept += 1
            temp *= self.change_temp
            max_angle *= self.change_angle
            if unaccept > 2000 or max_angle < 1e-7:
                break
        return e_old


__all__ = ["jacobi_rotation", "DociRotationModel", "OrbitalOptimization", "LocalMinimizer", "SimulatedAnnealing",
           "CF_DOCI", "CF_FCI"]
