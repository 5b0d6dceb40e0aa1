"""Brute-force second quantisation on tiny active spaces (test infrastructure).

Pins the oracle's Slater-Condon elements, phases and RDM conventions without
trusting any formula from the restatement: operators act on spin-orbital
occupation bitstrings with the Jordan-Wigner sign, spin-orbitals ordered alpha
0..L-1 then beta L..2L-1, which is the reference's convention that alpha
operators stand to the left of beta ones (CIMethod.cpp:409-457 count only
same-spin electrons for the phase).

  H = sum h_pq a+_ps a_qs + 1/2 sum V_pqrs a+_ps a+_qt a_st a_rs   (physics V)
  RDM  Gamma^{st}(ij,kl) = <a+_is a+_jt a_lt a_ks>   (CIDens.h:32-38)
"""
from __future__ import annotations

import itertools

import numpy as np


def _ann(state: int, p: int):
    if not (state >> p) & 1:
        return None, 0
    sign = -1 if bin(state & ((1 << p) - 1)).count("1") & 1 else 1
    return state ^ (1 << p), sign


def _cre(state: int, p: int):
    if (state >> p) & 1:
        return None, 0
    sign = -1 if bin(state & ((1 << p) - 1)).count("1") & 1 else 1
    return state | (1 << p), sign


def _apply(ops, state):
    """ops: list of (kind, p) applied right-to-left. kind 'c' or 'a'."""
    sign = 1
    for kind, p in reversed(ops):
        state, s = (_cre if kind == "c" else _ann)(state, p)
        if state is None:
            return None, 0
        sign *= s
    return state, sign


def strings(L: int, n: int):
    """All n-of-L strings in increasing integer order (= Gosper order)."""
    out = []
    for occ in itertools.combinations(range(L), n):
        out.append(sum(1 << o for o in occ))
    return sorted(out)


def basis(L: int, na: int, nb: int):
    A = strings(L, na)
    B = strings(L, nb)
    dets = [(a, b) for a in A for b in B]  # alpha slow, beta fast
    return dets


def spin_state(L, a, b):
    return a | (b << L)


def hamiltonian(h, V, na, nb, dets=None):
    L = h.shape[0]
    if dets is None:
        dets = basis(L, na, nb)
    index = {spin_state(L, a, b): i for i, (a, b) in enumerate(dets)}
    n = len(dets)
    H = np.zeros((n, n))
    so = lambda p, s: p + s * L  # noqa: E731
    for j, (a, b) in enumerate(dets):
        st = spin_state(L, a, b)
        for s in (0, 1):
            for p in range(L):
                for q in range(L):
                    if h[p, q] == 0.0:
                        continue
                    r, sg = _apply([("c", so(p, s)), ("a", so(q, s))], st)
                    if r is not None and r in index:
                        H[index[r], j] += sg * h[p, q]
        for s in (0, 1):
            for t in (0, 1):
                for p, q, r_, s_ in itertools.product(range(L), repeat=4):
                    v = V[p, q, r_, s_]
                    if v == 0.0:
                        continue
                    res, sg = _apply([("c", so(p, s)), ("c", so(q, t)), ("a", so(s_, t)), ("a", so(r_, s))], st)
                    if res is not None and res in index:
                        H[index[res], j] += 0.5 * sg * v
    return H


def rdms(c, L, na, nb, dets=None):
    """Dense gamma[s][i][j] and Gamma[b][i][j][k][l], b: 0 aa, 1 ab, 2 bb."""
    if dets is None:
        dets = basis(L, na, nb)
    index = {spin_state(L, a, b): i for i, (a, b) in enumerate(dets)}
    so = lambda p, s: p + s * L  # noqa: E731
    g = np.zeros((2, L, L))
    G = np.zeros((3, L, L, L, L))
    for j, (a, b) in enumerate(dets):
        cj = c[j]
        if cj == 0.0:
            continue
        st = spin_state(L, a, b)
        for s in (0, 1):
            for p in range(L):
                for q in range(L):
                    r, sg = _apply([("c", so(p, s)), ("a", so(q, s))], st)
                    if r is not None and r in index:
                        g[s, p, q] += sg * c[index[r]] * cj
        for blk, (s, t) in enumerate(((0, 0), (0, 1), (1, 1))):
            for i, jj, k, l in itertools.product(range(L), repeat=4):
                res, sg = _apply([("c", so(i, s)), ("c", so(jj, t)), ("a", so(l, t)), ("a", so(k, s))], st)
                if res is not None and res in index:
                    G[blk, i, jj, k, l] += sg * c[index[res]] * cj
    return g, G
