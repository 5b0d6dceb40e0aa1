"""The reference's HDF5 files through libciflow_h5.so (host only, C ABI in
include/ciflow_amd_h5.h).

* read_hamiltonian / save_hamiltonian: Hamiltonian::read / Hamiltonian::save
  (Hamiltonian.cpp:322-495) with the symmetry-packed TwoIndex / FourIndex
  blocks (TwoIndex.cpp:82-202, FourIndex.cpp:317-433), as dense h[L,L] and
  physics V[L,L,L,L];
* read_unitary / save_unitary: UnitaryMatrix::loadU / saveU
  (UnitaryMatrix.cpp:469-526);
* write_crs / read_crs: SparseMatrix_CRS::write_to_file / read_from_file
  (SparseMatrix_CRS.cpp:484-661), the upper CRS with 32-bit row pointers.

The library is built by ciflow_amd.build.build_h5 where an HDF5 C library is
installed; available() says whether it was.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libciflow_h5.so")

_lib = None
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)


class H5Error(RuntimeError):
    pass


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not available():
        raise ImportError(f"{LIB_PATH} not built: run python -m ciflow_amd.build (needs the HDF5 C library)")
    L = C.CDLL(LIB_PATH)
    L.cf_h5_last_error.restype = C.c_char_p
    L.cf_h5_last_error.argtypes = []
    L.cf_h5_read_hamiltonian.restype = C.c_int
    L.cf_h5_read_hamiltonian.argtypes = [C.c_char_p, ip, ip, ip, dp, ip, ip, dp, dp, dp, ip]
    L.cf_h5_save_hamiltonian.restype = C.c_int
    L.cf_h5_save_hamiltonian.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, ip, dp, dp, dp]
    L.cf_h5_read_unitary.restype = C.c_int
    L.cf_h5_read_unitary.argtypes = [C.c_char_p, C.c_int, C.c_int, ip, dp]
    L.cf_h5_save_unitary.restype = C.c_int
    L.cf_h5_save_unitary.argtypes = [C.c_char_p, C.c_int, C.c_int, ip, dp]
    L.cf_h5_write_crs.restype = C.c_int
    L.cf_h5_write_crs.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_uint64, C.c_uint64, u32p, u32p, dp]
    L.cf_h5_read_crs.restype = C.c_int
    L.cf_h5_read_crs.argtypes = [C.c_char_p, C.c_char_p, u64p, u64p, u32p, u32p, dp]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise H5Error(f"[{rc}] {lib().cf_h5_last_error().decode(errors='replace')}")


def _d(a):
    return None if a is None else a.ctypes.data_as(dp)


def read_hamiltonian(path: str) -> dict:
    """Hamiltonian::read: {L, nalpha, nbeta, econst, ngroup, irreps, h, V,
    overlap (None when the file holds none)}."""
    L, nu, nd, ng, ov = C.c_int(), C.c_int(), C.c_int(), C.c_int(), C.c_int()
    ec = C.c_double()
    b = os.fsencode(path)
    _check(lib().cf_h5_read_hamiltonian(b, C.byref(L), C.byref(nu), C.byref(nd), C.byref(ec), C.byref(ng), None,
                                        None, None, None, C.byref(ov)))
    n = L.value
    h = np.zeros((n, n))
    V = np.zeros((n, n, n, n))
    S = np.zeros((n, n)) if ov.value else None
    irr = np.zeros(n, dtype=np.int32)
    _check(lib().cf_h5_read_hamiltonian(b, C.byref(L), None, None, None, None, irr.ctypes.data_as(ip), _d(h), _d(V),
                                        _d(S), None))
    return {"L": n, "nalpha": nu.value, "nbeta": nd.value, "econst": ec.value, "ngroup": ng.value,
            "irreps": irr.astype(int), "h": h, "V": V, "overlap": S}


def save_hamiltonian(path: str, h, V, nalpha: int, nbeta: int, econst: float = 0.0, ngroup: int = 0,
                     irreps=None, overlap=None) -> None:
    """Hamiltonian::save to exactly `path`."""
    h = np.ascontiguousarray(h, dtype=np.float64)
    V = np.ascontiguousarray(V, dtype=np.float64)
    n = h.shape[0]
    if h.shape != (n, n) or V.shape != (n,) * 4:
        raise ValueError("h must be [L,L] and V [L,L,L,L]")
    irr = None if irreps is None else np.ascontiguousarray(irreps, dtype=np.int32)
    if irr is not None and irr.shape != (n,):
        raise ValueError("irreps must have L entries")
    S = None if overlap is None else np.ascontiguousarray(overlap, dtype=np.float64)
    if S is not None and S.shape != (n, n):
        raise ValueError("overlap must be [L,L]")
    _check(lib().cf_h5_save_hamiltonian(os.fsencode(path), n, int(nalpha), int(nbeta), float(econst), int(ngroup),
                                        None if irr is None else irr.ctypes.data_as(ip), _d(h), _d(V), _d(S)))


def read_unitary(path: str, irreps, ngroup: int) -> np.ndarray:
    """UnitaryMatrix::loadU: dense U[L,L], new orbitals in rows, the U that
    Hamiltonian.rotate applies."""
    irr = np.ascontiguousarray(irreps, dtype=np.int32)
    n = len(irr)
    U = np.zeros((n, n))
    _check(lib().cf_h5_read_unitary(os.fsencode(path), n, int(ngroup), irr.ctypes.data_as(ip), _d(U)))
    return U


def save_unitary(path: str, U, irreps, ngroup: int) -> None:
    """UnitaryMatrix::saveU of a dense, irrep-block-diagonal U[L,L]."""
    U = np.ascontiguousarray(U, dtype=np.float64)
    irr = np.ascontiguousarray(irreps, dtype=np.int32)
    if U.shape != (len(irr), len(irr)):
        raise ValueError("U must be [L,L] with L the number of orbital irreps")
    _check(lib().cf_h5_save_unitary(os.fsencode(path), len(irr), int(ngroup), irr.ctypes.data_as(ip), _d(U)))


def write_crs(path: str, name: str, row, col, data, append: bool = False) -> None:
    """SparseMatrix_CRS::write_to_file: group `name` holding row[n+1],
    col[nnz], data[nnz]."""
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    data = np.ascontiguousarray(data, dtype=np.float64)
    if row.ndim != 1 or len(row) < 1 or col.shape != data.shape:
        raise ValueError("row[n+1], col[nnz] and data[nnz] expected")
    _check(lib().cf_h5_write_crs(os.fsencode(path), name.encode(), int(bool(append)), len(row) - 1, len(data),
                                 row.ctypes.data_as(u32p), col.ctypes.data_as(u32p), _d(data)))


def read_crs(path: str, name: str):
    """SparseMatrix_CRS::read_from_file: (row, col, data) as uint32, uint32, float64."""
    n, nnz = C.c_uint64(), C.c_uint64()
    b = os.fsencode(path)
    _check(lib().cf_h5_read_crs(b, name.encode(), C.byref(n), C.byref(nnz), None, None, None))
    row = np.zeros(n.value + 1, dtype=np.uint32)
    col = np.zeros(max(nnz.value, 1), dtype=np.uint32)
    data = np.zeros(max(nnz.value, 1))
    _check(lib().cf_h5_read_crs(b, name.encode(), C.byref(n), C.byref(nnz), row.ctypes.data_as(u32p),
                                col.ctypes.data_as(u32p), _d(data)))
    return row, col[:nnz.value], data[:nnz.value]
