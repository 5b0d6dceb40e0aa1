"""ctypes front end of the CPU oracle (test infrastructure only).

The oracle restates the reference CIFlow algorithms on the CPU
(``ciflow_oracle.cpp``).  It is loaded only by ``tests/``, by
``__graft_entry__.smoke()`` and by the CPU-baseline leg of ``bench.py``, always
as the checker or as the timed CPU baseline, never as the thing measured on
the GPU.  The GPU library (``ciflow_amd``) does not import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64 = C.c_uint64
dp = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
ip = C.POINTER(C.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.or_gen_integrals.argtypes = [C.c_int, u64, dp, dp]
        L.or_first_string.argtypes = [C.c_int]
        L.or_first_string.restype = u64
        L.or_next_string.argtypes = [u64]
        L.or_next_string.restype = u64
        L.or_rank.argtypes = [C.c_int, C.c_int, u64]
        L.or_rank.restype = C.c_uint32
        L.or_binom.argtypes = [C.c_int, C.c_int]
        L.or_binom.restype = u64
        L.or_enumerate.argtypes = [C.c_int, C.c_int, u64p]
        L.or_build_crs.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, C.c_int, C.c_int,
                                   C.c_int]
        L.or_build_crs.restype = C.c_void_p
        L.or_crs_nrows.argtypes = [C.c_void_p]
        L.or_crs_nrows.restype = u64
        L.or_crs_nnz.argtypes = [C.c_void_p]
        L.or_crs_nnz.restype = u64
        L.or_crs_copy.argtypes = [C.c_void_p, u64p, u32p, dp]
        L.or_crs_free.argtypes = [C.c_void_p]
        L.or_crs_full_from_upper.argtypes = [u64, u64p, u32p, dp]
        L.or_crs_full_from_upper.restype = C.c_void_p
        L.or_ham_element.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, u64, u64, u64, u64]
        L.or_ham_element.restype = C.c_double
        L.or_mvprod_sym.argtypes = [u64, u64p, u32p, dp, dp, dp]
        L.or_mvprod_full.argtypes = [u64, u64p, u32p, dp, dp, dp, C.c_int]
        L.or_lanczos.argtypes = [u64, u64p, u32p, dp, C.c_int, C.c_double, C.c_int, C.c_int, dp, dp, ip, ip]
        L.or_lanczos.restype = C.c_int
        L.or_rdm_fci.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp]
        L.or_rdm_doci.argtypes = [C.c_int, C.c_int, dp, dp, dp]
        L.or_get_one_rdm.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_int]
        L.or_get_one_rdm.restype = C.c_double
        L.or_get_two_rdm.argtypes = [C.c_int, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_get_two_rdm.restype = C.c_double
        L.or_dens_energy.argtypes = [C.c_int, dp, dp, C.c_double, dp, dp]
        L.or_dens_energy.restype = C.c_double
        L.or_num_threads.restype = C.c_int
        L.or_rotate_integrals.argtypes = [C.c_int, dp, dp, dp, dp, dp]
        L.or_read_dets.argtypes = [C.c_char_p, u64, u64p, u64p]
        L.or_read_dets.restype = C.c_int64
        L.or_build_crs_dets.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, u64, u64p, u64p, C.c_int]
        L.or_build_crs_dets.restype = C.c_void_p
        _lib = L
    return _lib


def _d(a: np.ndarray):
    return a.ctypes.data_as(dp)


def gen_integrals(L: int, seed: int):
    """Seeded synthetic integrals: h (L,L) and physics V (L,L,L,L)."""
    h = np.zeros((L, L))
    V = np.zeros((L, L, L, L))
    lib().or_gen_integrals(L, seed, _d(h), _d(V))
    return h, V


def binom(n: int, k: int) -> int:
    return int(lib().or_binom(n, k))


def enumerate_strings(L: int, n: int) -> np.ndarray:
    d = binom(L, n)
    out = np.zeros(d, dtype=np.uint64)
    lib().or_enumerate(L, n, out.ctypes.data_as(u64p))
    return out


def rank(L: int, n: int, s: int) -> int:
    return int(lib().or_rank(L, n, s))


class CRS:
    """Upper-triangle CRS in the reference layout (row, col, data)."""

    def __init__(self, row, col, data):
        self.row = row
        self.col = col
        self.data = data

    @property
    def n(self):
        return len(self.row) - 1

    @property
    def nnz(self):
        return len(self.data)


def _take_crs(h) -> CRS:
    L = lib()
    nr = int(L.or_crs_nrows(h))
    nz = int(L.or_crs_nnz(h))
    row = np.zeros(nr + 1, dtype=np.uint64)
    col = np.zeros(nz, dtype=np.uint32)
    val = np.zeros(nz)
    L.or_crs_copy(h, row.ctypes.data_as(u64p), col.ctypes.data_as(u32p), _d(val))
    L.or_crs_free(h)
    return CRS(row, col, val)


def build_crs(kind: str, h, V, nup: int, ndown: int, start: int = 0, end: int | None = None,
              zero: float = 1e-14, threads: int = 0) -> CRS:
    """FCI/DOCI::construct_CI_matrix over lines [start, end) via build_parallel."""
    L = h.shape[0]
    k = 0 if kind == "FCI" else 1
    if end is None:
        end = binom(L, nup)
    h = np.ascontiguousarray(h, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    if threads <= 0:
        threads = int(lib().or_num_threads())
    hh = lib().or_build_crs(k, L, nup, ndown, _d(h), _d(V), zero, start, end, threads)
    return _take_crs(hh)


def read_dets(path: str):
    """Permutator_File::read_file: (up, down) uint64 arrays in file order."""
    n = int(lib().or_read_dets(path.encode(), 0, None, None))
    if n < 0:
        raise ValueError(f"cannot read determinants from {path} ({n})")
    up = np.zeros(n, dtype=np.uint64)
    down = np.zeros(n, dtype=np.uint64)
    lib().or_read_dets(path.encode(), n, up.ctypes.data_as(u64p), down.ctypes.data_as(u64p))
    return up, down


def build_crs_dets(h, V, nup: int, ndown: int, up, down, zero: float = 1e-14, threads: int = 0) -> CRS:
    """FCI_File::construct_CI_matrix over a determinant list via build_parallel."""
    L = h.shape[0]
    h = np.ascontiguousarray(h, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    up = np.ascontiguousarray(up, dtype=np.uint64)
    down = np.ascontiguousarray(down, dtype=np.uint64)
    if threads <= 0:
        threads = int(lib().or_num_threads())
    hh = lib().or_build_crs_dets(L, nup, ndown, _d(h), _d(V), zero, len(up), up.ctypes.data_as(u64p),
                                 down.ctypes.data_as(u64p), threads)
    return _take_crs(hh)


def full_from_upper(m: CRS) -> CRS:
    hh = lib().or_crs_full_from_upper(m.n, m.row.ctypes.data_as(u64p), m.col.ctypes.data_as(u32p), _d(m.data))
    return _take_crs(hh)


def ham_element(h, V, nup, ndown, up1, down1, up2, down2) -> float:
    L = h.shape[0]
    return float(lib().or_ham_element(L, nup, ndown, _d(h), _d(V), up1, down1, up2, down2))


def mvprod_sym(m: CRS, x: np.ndarray) -> np.ndarray:
    y = np.zeros(m.n)
    lib().or_mvprod_sym(m.n, m.row.ctypes.data_as(u64p), m.col.ctypes.data_as(u32p), _d(m.data), _d(x), _d(y))
    return y


def mvprod_full(m: CRS, x: np.ndarray, threads: int = 0) -> np.ndarray:
    y = np.zeros(m.n)
    if threads <= 0:
        threads = int(lib().or_num_threads())
    lib().or_mvprod_full(m.n, m.row.ctypes.data_as(u64p), m.col.ctypes.data_as(u32p), _d(m.data), _d(x), _d(y),
                         threads)
    return y


def dense_from_upper(m: CRS) -> np.ndarray:
    n = m.n
    H = np.zeros((n, n))
    for i in range(n):
        a, b = int(m.row[i]), int(m.row[i + 1])
        cols = m.col[a:b].astype(np.int64)
        H[i, cols] = m.data[a:b]
        H[cols, i] = m.data[a:b]
    return H


def lanczos(full: CRS, ncv: int = 42, tol: float = 1e-12, maxrestart: int | None = None, threads: int = 0):
    """Thick-restart Lanczos on a full-row CRS. Returns (E, vec, info)."""
    n = full.n
    if maxrestart is None:
        maxrestart = 3 * n
    if threads <= 0:
        threads = int(lib().or_num_threads())
    ev = C.c_double()
    vec = np.zeros(n)
    nmv = C.c_int()
    nrs = C.c_int()
    st = lib().or_lanczos(n, full.row.ctypes.data_as(u64p), full.col.ctypes.data_as(u32p), _d(full.data), ncv,
                          tol, maxrestart, threads, C.byref(ev), _d(vec), C.byref(nmv), C.byref(nrs))
    return ev.value, vec, {"status": st, "n_matvec": nmv.value, "n_restart": nrs.value}


def rdm_sizes(L: int):
    return 2 * L * (L + 1) // 2, 3 * (L * L) * (L * L + 1) // 2


def rdm_fci(L, nup, ndown, c, two=True):
    so, st = rdm_sizes(L)
    one = np.zeros(so)
    tw = np.zeros(st) if two else None
    c = np.ascontiguousarray(c, dtype=np.float64)
    lib().or_rdm_fci(L, nup, ndown, _d(c), _d(one), _d(tw) if two else None)
    return one, tw


def rdm_doci(L, npair, c, two=True):
    so, st = rdm_sizes(L)
    one = np.zeros(so)
    tw = np.zeros(st) if two else None
    c = np.ascontiguousarray(c, dtype=np.float64)
    lib().or_rdm_doci(L, npair, _d(c), _d(one), _d(tw) if two else None)
    return one, tw


def get_one_rdm(L, one, spin, a, b):
    return float(lib().or_get_one_rdm(L, _d(one), spin, a, b))


def get_two_rdm(L, two, spin, a, b, c, d):
    return float(lib().or_get_two_rdm(L, _d(two), spin, a, b, c, d))


def dens_energy(h, V, econst, one, two):
    L = h.shape[0]
    h = np.ascontiguousarray(h)
    V = np.ascontiguousarray(V)
    return float(lib().or_dens_energy(L, _d(h), _d(V), econst, _d(one), _d(two)))


def num_threads() -> int:
    return int(lib().or_num_threads())


def rotate_integrals(U, h, V):
    """OrbitalTransform::fillHamCI restated (quarter transforms, C1)."""
    L = h.shape[0]
    U = np.ascontiguousarray(U, dtype=np.float64)
    h = np.ascontiguousarray(h, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    ho = np.zeros_like(h)
    Vo = np.zeros_like(V)
    lib().or_rotate_integrals(L, _d(U), _d(h), _d(V), _d(ho), _d(Vo))
    return ho, Vo
