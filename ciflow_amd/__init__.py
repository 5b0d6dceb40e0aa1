"""ciflow_amd -- MI355X (gfx950) CI ground-state hot path of CIFlow.

Python front end over the C ABI in include/ciflow_amd.h (libciflow_amd.so,
hand-written HIP kernels).  The classes mirror the reference's interface:

    Hamiltonian            Hamiltonian.h:80-136   (getTmat/getVmat/getEconst/getL/getnup/getndown)
    FCI, DOCI              CIMethod.h:135-162     (build_ci, solve, get_ci_energy, get_eigvec,
                                                   construct_density, get_one_rdm/get_two_rdm,
                                                   get_dens_energy)
    SparseMatrix_CRS_Sym   SparseMatrix_CRS.h:120-130 (mvprod, arpackDiagonalize)

There is no CPU fallback: importing works everywhere, but every compute call
needs the built library and a GPU, and raises if either is missing.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libciflow_amd.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ciflow_amd.h")

CF_OK, CF_ERR_ARG, CF_ERR_HIP, CF_ERR_NOMEM, CF_ERR_NOCONV, CF_ERR_UNSUPPORTED, CF_ERR_STATE = range(7)
CF_FCI, CF_DOCI, CF_DETS = 0, 1, 2
CF_F64, CF_F32 = 0, 1
CF_PREC_AUTO = 2  # served requests: FP32 values only where the operator stores them
CF_LANCZOS, CF_DAVIDSON, CF_SOLVER_AUTO = 0, 1, 2

_lib = None
_exiting = False


def _mark_exit():
    global _exiting
    _exiting = True


atexit.register(_mark_exit)


class _Handle:
    """Owns one native handle; keeps its parent handle alive so native objects
    are destroyed child-first.  Destroy calls are skipped at interpreter exit
    (the process tears the device context down)."""

    def __init__(self, destroy: str, parent=None):
        self.ptr = C.c_void_p()
        self._destroy = destroy
        self.parent = parent

    def __del__(self):
        if self.ptr and not _exiting and _lib is not None:
            try:
                getattr(_lib, self._destroy)(self.ptr)
            except Exception:
                pass
        self.ptr = C.c_void_p()


dp = C.POINTER(C.c_double)
u64p = C.POINTER(C.c_uint64)
u32p = C.POINTER(C.c_uint32)
vp = C.c_void_p


class CIFlowError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ciflow_amd error {code}: {msg}")
        self.code = code


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double, C.c_double)


class TuneEntry(C.Structure):
    """cf_tune_entry: one timed SpMV launch shape."""
    _fields_ = [("kernel", C.c_int), ("p1", C.c_int), ("p2", C.c_int), ("ms", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("ncv", C.c_int), ("maxrestart", C.c_int), ("tol", C.c_double), ("precision", C.c_int),
                ("tol_f32", C.c_double), ("seed", C.c_uint64), ("check_every", C.c_int),
                ("tol_energy", C.c_double), ("progress", PROGRESS_FN), ("progress_user", C.c_void_p),
                ("method", C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [("eigval", C.c_double), ("energy", C.c_double), ("resid", C.c_double), ("n_matvec", C.c_int),
                ("n_matvec_f32", C.c_int), ("n_restart", C.c_int), ("status", C.c_int), ("ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("kind", C.c_int), ("L", C.c_int), ("nup", C.c_int), ("ndown", C.c_int),
                ("h", dp), ("V", dp), ("U", dp), ("econst", C.c_double), ("precision", C.c_int), ("ncv", C.c_int),
                ("tol", C.c_double), ("check_every", C.c_int), ("tol_energy", C.c_double), ("direct", C.c_int),
                ("method", C.c_int)]


class Result(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int), ("solve_status", C.c_int), ("n", C.c_uint64),
                ("energy", C.c_double), ("eigval", C.c_double), ("resid", C.c_double), ("n_matvec", C.c_int),
                ("n_matvec_f32", C.c_int), ("t_submit", C.c_double), ("t_start", C.c_double),
                ("t_end", C.c_double), ("error", C.c_char * 160)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["error"] = self.error.decode(errors="replace")
        return d


def header_symbols() -> list[str]:
    """Entry points declared in include/ciflow_amd.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(cf_\w+)\s*\(", txt, re.M)))


def lib():
    """Load libciflow_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run python -m ciflow_amd.build")
    L = C.CDLL(LIB_PATH)
    sig = {
        "cf_version": (C.c_int, []),
        "cf_last_error": (C.c_char_p, []),
        "cf_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "cf_ctx_destroy": (C.c_int, [vp]),
        "cf_ctx_synchronize": (C.c_int, [vp]),
        "cf_ctx_enable_timing": (C.c_int, [vp, C.c_int]),
        "cf_ctx_kernel_time": (C.c_int, [vp, C.c_char_p, C.POINTER(C.c_int64), dp]),
        "cf_ctx_reset_timing": (C.c_int, [vp]),
        "cf_ctx_stream_copy": (C.c_int, [vp, C.c_uint64, C.c_int]),
        "cf_gen_integrals": (C.c_int, [C.c_int, C.c_uint64, dp, dp]),
        "cf_ham_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, dp, C.c_double, C.POINTER(vp)]),
        "cf_ham_destroy": (C.c_int, [vp]),
        "cf_space_create": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "cf_space_create_dets": (C.c_int, [vp, C.c_uint64, u64p, u64p, C.POINTER(vp)]),
        "cf_read_determinants": (C.c_int, [C.c_char_p, C.c_uint64, u64p, u64p, C.POINTER(C.c_uint64)]),
        "cf_read_integrals": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                        C.POINTER(C.c_double), dp, dp, C.POINTER(C.c_int)]),
        "cf_space_destroy": (C.c_int, [vp]),
        "cf_space_set_ham": (C.c_int, [vp, vp]),
        "cf_space_dims": (C.c_int, [vp, u64p, u64p, u64p]),
        "cf_strings_unrank": (C.c_int, [vp, C.c_int, C.c_int, C.c_uint64, C.c_uint64, u64p]),
        "cf_strings_rank": (C.c_int, [vp, C.c_int, C.c_int, u64p, C.c_uint64, u32p]),
        "cf_crs_build": (C.c_int, [vp, C.c_int, C.c_double, C.POINTER(vp)]),
        "cf_crs_destroy": (C.c_int, [vp]),
        "cf_crs_info": (C.c_int, [vp, u64p, u64p, u64p]),
        "cf_crs_count_upper": (C.c_int, [vp, C.c_uint64, C.c_uint64, u64p]),
        "cf_crs_export_upper": (C.c_int, [vp, C.c_uint64, C.c_uint64, u64p, u32p, dp]),
        "cf_crs_export_upper_u32": (C.c_int, [vp, C.c_uint64, C.c_uint64, u32p, u32p, dp]),
        "cf_crs_count_full": (C.c_int, [vp, C.c_uint64, C.c_uint64, u64p]),
        "cf_crs_export_full": (C.c_int, [vp, C.c_uint64, C.c_uint64, u64p, u32p, dp]),
        "cf_mvprod": (C.c_int, [vp, dp, dp]),
        "cf_mvprod_dev": (C.c_int, [vp, C.c_int, vp, vp]),
        "cf_mvprod_prec": (C.c_int, [vp, C.c_int, dp, dp]),
        "cf_crs_spmv_bytes": (C.c_int, [vp, C.c_int, u64p, C.POINTER(C.c_int)]),
        "cf_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
        "cf_solve": (C.c_int, [vp, C.POINTER(SolveOpts), dp, C.POINTER(SolveInfo)]),
        "cf_rci_create": (C.c_int, [C.c_uint64, C.c_int, C.c_double, C.c_int, C.c_uint64, C.POINTER(vp)]),
        "cf_rci_step": (C.c_int, [vp, C.POINTER(C.c_int), dp, C.POINTER(C.c_int)]),
        "cf_rci_result": (C.c_int, [vp, dp, dp, C.POINTER(C.c_int)]),
        "cf_rci_destroy": (C.c_int, [vp]),
        "cf_rdm": (C.c_int, [vp, dp, dp, dp]),
        "cf_crs_set_spmv_kernel": (C.c_int, [vp, C.c_int, C.c_int]),
        "cf_crs_set_spmv_shape": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "cf_crs_autotune": (C.c_int, [vp, C.c_int, C.c_int, C.POINTER(TuneEntry), C.c_int, C.POINTER(C.c_int)]),
        "cf_rotate_integrals": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp, dp, dp]),
        "cf_rotate_integrals_noalloc": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]),
        "cf_ham_rotate": (C.c_int, [vp, dp, C.POINTER(vp)]),
        "cf_ham_get": (C.c_int, [vp, dp, dp]),
        "cf_expm_skew": (C.c_int, [C.c_int, dp, dp]),
        "cf_gen_rotation": (C.c_int, [C.c_int, C.c_uint64, C.c_double, dp]),
        "cf_sigma_create": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
        "cf_dets_sigma_create": (C.c_int, [vp, C.c_uint64, C.POINTER(vp)]),
        "cf_sigma_destroy": (C.c_int, [vp]),
        "cf_sigma_set_threads": (C.c_int, [vp, C.c_int]),
        "cf_sigma_set_kernel": (C.c_int, [vp, C.c_int]),
        "cf_sigma_apply_dev": (C.c_int, [vp, C.c_int, vp, vp]),
        "cf_sigma_apply": (C.c_int, [vp, C.c_int, dp, dp]),
        "cf_solve_direct": (C.c_int, [vp, C.POINTER(SolveOpts), dp, C.POINTER(SolveInfo)]),
        "cf_rdm1_direct": (C.c_int, [vp, dp, dp]),
        "cf_space_eigvec": (C.c_int, [vp, dp]),
        "cf_rdm_dense": (C.c_int, [vp, dp, dp, dp]),
        "cf_server_create": (C.c_int, [C.c_int, C.c_int, C.c_double, C.POINTER(vp)]),
        "cf_server_submit": (C.c_int, [vp, C.POINTER(Request)]),
        "cf_server_poll": (C.c_int, [vp, C.POINTER(Result), C.c_int, C.POINTER(C.c_int)]),
        "cf_server_drain": (C.c_int, [vp, C.c_double]),
        "cf_server_stats": (C.c_int, [vp, u64p, u64p, u64p, u64p]),
        "cf_server_destroy": (C.c_int, [vp]),
        "cf_request_footprint": (C.c_int, [C.POINTER(Request), u64p]),
        "cf_solve_doci_batch": (C.c_int, [vp, C.c_int, C.POINTER(Request), C.POINTER(Result)]),
        "cf_solve_batch": (C.c_int, [vp, C.c_int, C.POINTER(Request), C.POINTER(Result)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(code: int):
    if code != CF_OK:
        msg = lib().cf_last_error()
        raise CIFlowError(code, msg.decode() if msg else "")


def _d(a: np.ndarray):
    return a.ctypes.data_as(dp)


def gen_integrals(L: int, seed: int):
    """Seeded synthetic integrals (h[L,L], physics V[L,L,L,L])."""
    h = np.zeros((L, L))
    V = np.zeros((L, L, L, L))
    _check(lib().cf_gen_integrals(L, seed, _d(h), _d(V)))
    return h, V


def expm_skew(K) -> np.ndarray:
    """U = exp(K) for antisymmetric K (UnitaryMatrix::updateUnitary)."""
    K = np.ascontiguousarray(K, dtype=np.float64)
    U = np.zeros_like(K)
    _check(lib().cf_expm_skew(K.shape[0], _d(K), _d(U)))
    return U


def gen_rotation(L: int, seed: int, sigma: float = 0.05) -> np.ndarray:
    K = np.zeros((L, L))
    _check(lib().cf_gen_rotation(L, seed, sigma, _d(K)))
    return K


def rotate_integrals(h, V, U, ctx: "Context | None" = None):
    """Batched OrbitalTransform::fillHamCI on the GPU: h[b,L,L], V[b,L,L,L,L], U[b,L,L]
    (a single Hamiltonian without the batch axis is accepted too)."""
    ctx = ctx or default_context()
    single = np.ndim(h) == 2
    h = np.ascontiguousarray(h[None] if single else h, dtype=np.float64)
    V = np.ascontiguousarray(V[None] if single else V, dtype=np.float64)
    U = np.ascontiguousarray(U[None] if single else U, dtype=np.float64)
    B, L = h.shape[0], h.shape[1]
    ho = np.zeros_like(h)
    Vo = np.zeros_like(V)
    _check(lib().cf_rotate_integrals(ctx.ptr, L, B, _d(h), _d(V), _d(U), _d(ho), _d(Vo)))
    return (ho[0], Vo[0]) if single else (ho, Vo)


class Context:
    """One GPU + one stream (cf_ctx)."""

    def __init__(self, device: int = 0):
        self.h = _Handle("cf_ctx_destroy")
        _check(lib().cf_ctx_create(device, C.byref(self.h.ptr)))

    @property
    def ptr(self):
        return self.h.ptr

    def synchronize(self):
        _check(lib().cf_ctx_synchronize(self.ptr))

    def enable_timing(self, on: bool = True):
        _check(lib().cf_ctx_enable_timing(self.ptr, int(on)))

    def reset_timing(self):
        _check(lib().cf_ctx_reset_timing(self.ptr))

    def kernel_time(self, family: str):
        n = C.c_int64()
        ms = C.c_double()
        _check(lib().cf_ctx_kernel_time(self.ptr, family.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def unrank(self, L: int, nel: int, first: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=np.uint64)
        _check(lib().cf_strings_unrank(self.ptr, L, nel, first, count, out.ctypes.data_as(u64p)))
        return out

    def rank(self, L: int, nel: int, strings) -> np.ndarray:
        s = np.ascontiguousarray(strings, dtype=np.uint64)
        out = np.zeros(len(s), dtype=np.uint32)
        _check(lib().cf_strings_rank(self.ptr, L, nel, s.ctypes.data_as(u64p), len(s), out.ctypes.data_as(u32p)))
        return out


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


class Hamiltonian:
    """Integrals on the GPU (Hamiltonian.h:80-136): h[i,j] = getTmat(i,j),
    V[i,j,k,l] = getVmat(i,j,k,l) (physics), econst = getEconst()."""

    def __init__(self, h, V, nup: int, ndown: int, econst: float = 0.0, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.h = np.ascontiguousarray(h, dtype=np.float64)
        self.V = np.ascontiguousarray(V, dtype=np.float64)
        self.L = self.h.shape[0]
        assert self.h.shape == (self.L, self.L) and self.V.shape == (self.L,) * 4
        self.nup, self.ndown, self.econst = int(nup), int(ndown), float(econst)
        self.handle = _Handle("cf_ham_destroy", self.ctx.h)
        _check(lib().cf_ham_create(self.ctx.ptr, self.L, self.nup, self.ndown, _d(self.h), _d(self.V), self.econst,
                                   C.byref(self.handle.ptr)))

    @property
    def ptr(self):
        return self.handle.ptr

    @classmethod
    def _adopt(cls, handle, L, nup, ndown, econst, ctx):
        self = cls.__new__(cls)
        self.ctx = ctx
        self.handle = handle
        self.L, self.nup, self.ndown, self.econst = L, nup, ndown, econst
        self.h = np.zeros((L, L))
        self.V = np.zeros((L, L, L, L))
        _check(lib().cf_ham_get(handle.ptr, _d(self.h), _d(self.V)))
        return self

    def rotate(self, U) -> "Hamiltonian":
        """OrbitalTransform::fillHamCI on the GPU: a new Hamiltonian with
        h' = U h U^T, V'_abcd = sum U_ai U_bj U_ck U_dl V_ijkl."""
        U = np.ascontiguousarray(U, dtype=np.float64)
        hd = _Handle("cf_ham_destroy", self.ctx.h)
        _check(lib().cf_ham_rotate(self.ptr, _d(U), C.byref(hd.ptr)))
        return Hamiltonian._adopt(hd, self.L, self.nup, self.ndown, self.econst, self.ctx)

    @classmethod
    def from_file(cls, path: str, ctx: Context | None = None) -> "Hamiltonian":
        """Hamiltonian(filename) of the reference: read a psi4 integral dump."""
        d = read_integrals(path)
        ham = cls(d["h"], d["V"], d["nalpha"], d["nbeta"], d["econst"], ctx)
        ham.irreps = d["irreps"]
        return ham

    @classmethod
    def read(cls, path: str, ctx: Context | None = None) -> "Hamiltonian":
        """Hamiltonian::read (Hamiltonian.cpp:402-495): the HDF5 file of
        Hamiltonian::save, symmetry-packed integrals expanded to dense
        arrays (host library libciflow_h5.so)."""
        from . import h5

        d = h5.read_hamiltonian(path)
        ham = cls(d["h"], d["V"], d["nalpha"], d["nbeta"], d["econst"], ctx)
        ham.irreps, ham.ngroup, ham.overlap = d["irreps"], d["ngroup"], d["overlap"]
        return ham

    def save(self, path: str, ngroup: int | None = None, irreps=None, overlap=None) -> None:
        """Hamiltonian::save (Hamiltonian.cpp:322-400) to exactly `path`;
        the point group and orbital irreps default to those the Hamiltonian
        was read with (C1 otherwise)."""
        from . import h5

        irreps = getattr(self, "irreps", None) if irreps is None else irreps
        if ngroup is None:
            ngroup = getattr(self, "ngroup", 0 if irreps is None or not np.any(np.asarray(irreps)) else None)
            if ngroup is None:
                raise ValueError("orbital irreps given without their point group (ngroup 0-7)")
        overlap = getattr(self, "overlap", None) if overlap is None else overlap
        h5.save_hamiltonian(path, self.h, self.V, self.nup, self.ndown, self.econst, ngroup, irreps, overlap)

    @classmethod
    def synthetic(cls, L: int, nup: int, ndown: int, seed: int, ctx: Context | None = None):
        h, V = gen_integrals(L, seed)
        return cls(h, V, nup, ndown, 0.0, ctx)

    @classmethod
    def from_npz(cls, path: str, ctx: Context | None = None):
        d = np.load(path)
        return cls(d["h"], d["V"], int(d["nalpha"]), int(d["nbeta"]), float(d["econst"]), ctx)

    def getL(self):
        return self.L

    def getnup(self):
        return self.nup

    def getndown(self):
        return self.ndown

    def getEconst(self):
        return self.econst

    def getTmat(self, i, j):
        return float(self.h[i, j])

    def getVmat(self, i, j, k, l):
        return float(self.V[i, j, k, l])


class SparseMatrix_CRS_Sym:
    """GPU CRS Hamiltonian (full rows stored; the reference's upper-triangle
    layout is available through export_upper)."""

    def __init__(self, space: "_Handle", ctx: Context, value_prec: int = CF_F64, zero: float = 1e-14):
        self.ctx = ctx
        self.handle = _Handle("cf_crs_destroy", space)
        _check(lib().cf_crs_build(space.ptr, value_prec, zero, C.byref(self.handle.ptr)))
        self.value_prec = value_prec

    @property
    def ptr(self):
        return self.handle.ptr

    def info(self):
        n, nf, nu = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().cf_crs_info(self.ptr, C.byref(n), C.byref(nf), C.byref(nu)))
        return n.value, nf.value, nu.value

    def getn(self):
        n = C.c_uint64()
        _check(lib().cf_crs_info(self.ptr, C.byref(n), None, None))  # no CRS counts needed
        return n.value

    def datasize(self):
        """Stored upper-triangle entries, as the reference's data.size()."""
        return self.info()[2]

    def _export(self, upper: bool, r0: int, r1: int):
        cnt = C.c_uint64()
        f = lib().cf_crs_count_upper if upper else lib().cf_crs_count_full
        _check(f(self.ptr, r0, r1, C.byref(cnt)))
        row = np.zeros(r1 - r0 + 1, dtype=np.uint64)
        col = np.zeros(max(cnt.value, 1), dtype=np.uint32)
        val = np.zeros(max(cnt.value, 1))
        g = lib().cf_crs_export_upper if upper else lib().cf_crs_export_full
        _check(g(self.ptr, r0, r1, row.ctypes.data_as(u64p), col.ctypes.data_as(u32p), _d(val)))
        return row, col[:cnt.value], val[:cnt.value]

    def export_upper(self, r0: int = 0, r1: int | None = None):
        return self._export(True, r0, self.getn() if r1 is None else r1)

    def export_upper_u32(self, r0: int = 0, r1: int | None = None):
        """export_upper with the reference's 32-bit row pointer (SparseMatrix_CRS.h:111-112)."""
        r1 = self.getn() if r1 is None else r1
        cnt = C.c_uint64()
        _check(lib().cf_crs_count_upper(self.ptr, r0, r1, C.byref(cnt)))
        row = np.zeros(r1 - r0 + 1, dtype=np.uint32)
        col = np.zeros(max(cnt.value, 1), dtype=np.uint32)
        val = np.zeros(max(cnt.value, 1))
        _check(lib().cf_crs_export_upper_u32(self.ptr, r0, r1, row.ctypes.data_as(u32p), col.ctypes.data_as(u32p),
                                             _d(val)))
        return row, col[:cnt.value], val[:cnt.value]

    def export_full(self, r0: int = 0, r1: int | None = None):
        return self._export(False, r0, self.getn() if r1 is None else r1)

    def write_to_file(self, filename: str, name: str, append: bool = False) -> None:
        """SparseMatrix_CRS::write_to_file (SparseMatrix_CRS.cpp:484-571):
        the upper CRS as group `name` of an HDF5 file; read it back with
        ciflow_amd.h5.read_crs (SparseMatrix_CRS::read_from_file)."""
        from . import h5

        row, col, val = self.export_upper_u32()
        h5.write_crs(filename, name, row, col, val, append)

    def spmv_bytes(self, precision: int = CF_F64):
        """(kernel id 1-6, algorithmic HBM bytes per product) of the current choice."""
        b = C.c_uint64()
        k = C.c_int()
        _check(lib().cf_crs_spmv_bytes(self.ptr, precision, C.byref(b), C.byref(k)))
        return k.value, b.value

    def set_spmv_kernel(self, kind: int, tile_rows: int = 0):
        """0 automatic, 1 plain row gather, 2 FCI alpha-block staged, 3 vectorised
        row gather, 4 FCI position-major rows, 5 DOCI pair-hop columns
        (tile_rows = workgroup size for 4 and 5), 6 DOCI block-staged
        (tile_rows = low orbitals), 7 sliced rows: 64-row chunks with
        interleaved entries, a second copy of the matrix (tile_rows = waves per
        chunk, 1, 2, 4, 8 or 16)."""
        _check(lib().cf_crs_set_spmv_kernel(self.ptr, kind, tile_rows))

    def set_spmv_shape(self, precision: int, kernel: int, p1: int, p2: int = 0):
        """Launch shape for one value precision (see cf_crs_set_spmv_shape)."""
        _check(lib().cf_crs_set_spmv_shape(self.ptr, precision, kernel, p1, p2))

    def autotune(self, precision: int = CF_F64, reps: int = 3):
        """Times every supported launch shape, installs the fastest for this
        value precision and returns [(kernel, p1, p2, ms), ...]."""
        cap = 64
        tab = (TuneEntry * cap)()
        cnt = C.c_int()
        _check(lib().cf_crs_autotune(self.ptr, precision, reps, tab, cap, C.byref(cnt)))
        return [(t.kernel, t.p1, t.p2, t.ms) for t in tab[:min(cnt.value, cap)]]

    def mvprod(self, x: np.ndarray, precision: int = CF_F64) -> np.ndarray:
        """y = A x (SparseMatrix_CRS_Sym::mvprod); precision CF_F32 uses the FP32 values."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        if precision == CF_F64:
            _check(lib().cf_mvprod(self.ptr, _d(x), _d(y)))
        else:
            _check(lib().cf_mvprod_prec(self.ptr, precision, _d(x), _d(y)))
        return y

    def arpackDiagonalize(self, opts: SolveOpts | None = None, want_vec: bool = True):
        n = self.getn()
        if opts is None:
            opts = SolveOpts()
            lib().cf_solve_opts_default(C.byref(opts))
        vec = np.zeros(n) if want_vec else None
        info = SolveInfo()
        _check(lib().cf_solve(self.ptr, C.byref(opts), _d(vec) if want_vec else None, C.byref(info)))
        return info, vec


def solve_opts(**kw) -> SolveOpts:
    """cf_solve_opts with defaults; progress=callable(n_matvec, theta, resid) is
    wrapped as the C callback (kept alive on the returned object)."""
    o = SolveOpts()
    lib().cf_solve_opts_default(C.byref(o))
    prog = kw.pop("progress", None)
    for k, v in kw.items():
        setattr(o, k, v)
    if prog is not None:
        o._cb = PROGRESS_FN(lambda user, nmv, th, r: prog(nmv, th, r))
        o.progress = o._cb
    return o


class CIMethod:
    """Common part of FCI and DOCI (CIMethod.h:37-132)."""

    kind = CF_FCI

    def __init__(self, ham: Hamiltonian, value_prec: int = CF_F64, build: bool = True, zero: float = 1e-14):
        self.ham = ham
        self.ctx = ham.ctx
        self.space = _Handle("cf_space_destroy", ham.handle)
        self._create_space()
        n, da, db = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _check(lib().cf_space_dims(self.ptr, C.byref(n), C.byref(da), C.byref(db)))
        self.n, self.da, self.db = n.value, da.value, db.value
        self._zero = zero
        self._value_prec = value_prec
        self._mat = None
        self._eigval = None
        self._eigvec = None
        self._info = None
        self._one = None
        self._two = None
        if build:
            self.build_ci()

    def _create_space(self):
        _check(lib().cf_space_create(self.ham.ptr, self.kind, C.byref(self.space.ptr)))

    # CIMethod::build_ci -> build_parallel (CIMethod.cpp:489-561)
    def build_ci(self):
        self._mat = None
        self._mat = SparseMatrix_CRS_Sym(self.space, self.ctx, self._value_prec, self._zero)

    def set_sparsity(self, zero: float):
        self._zero = zero

    # CIMethod::set_ham (CIMethod.cpp:117-126)
    def set_ham(self, ham: Hamiltonian, build: bool = True):
        """New integrals for this method; the solution and densities are reset.
        A Hamiltonian with the same L, nalpha and nbeta on the same context
        keeps the space and its matrix resident: cf_space_set_ham refills the
        values in place (and the CRS arrays, pattern included, when they are
        read).  Any other Hamiltonian gets a new space and matrix."""
        if ham.ctx is self.ctx and (ham.L, ham.nup, ham.ndown) == (self.ham.L, self.ham.nup, self.ham.ndown):
            _check(lib().cf_space_set_ham(self.ptr, ham.ptr))
            self.space.parent = ham.handle
            self.ham = ham
            self._reset_solution()
            if build and self._mat is None:
                self.build_ci()
        else:
            self._new_space(ham, build)

    def _new_space(self, ham: Hamiltonian, build: bool):
        CIMethod.__init__(self, ham, value_prec=self._value_prec, build=build, zero=self._zero)

    def _reset_solution(self):
        self._eigval = self._eigvec = self._info = None
        self._one = self._two = None

    def get_dim(self):
        return self.n

    def gUpDim(self):
        return self.da

    def gDownDim(self):
        return self.db

    def get_mat(self) -> SparseMatrix_CRS_Sym:
        return self._mat

    # CIMethod::solve (CIMethod.cpp:241-264)
    def solve(self, precision: int = CF_F64, **kw):
        if self._mat is None:
            self.build_ci()
        opts = solve_opts(precision=precision, **kw)
        self._info, self._eigvec = self._mat.arpackDiagonalize(opts)
        self._eigval = self._info.eigval
        self._one = self._two = None
        return self._info

    def get_ci_energy(self):
        return self._eigval + self.ham.econst

    def get_eigvec(self):
        return self._eigvec

    # CIMethod::construct_density (CIMethod.cpp:161-171)
    def construct_density(self, trdm: bool = True):
        L = self.ham.L
        one = np.zeros(L * (L + 1))
        two = np.zeros(3 * (L * L) * (L * L + 1) // 2) if trdm else None
        c = np.ascontiguousarray(self._eigvec)
        _check(lib().cf_rdm(self.ptr, _d(c), _d(one), _d(two) if trdm else None))
        self._one, self._two = one, two
        return one, two

    def density_dense(self, c=None, two: bool = True):
        L = self.ham.L
        c = np.ascontiguousarray(self._eigvec if c is None else c, dtype=np.float64)
        g = np.zeros((2, L, L))
        G = np.zeros((3, L, L, L, L)) if two else None
        _check(lib().cf_rdm_dense(self.ptr, _d(c), _d(g), _d(G) if two else None))
        return g, G

    # CIDens accessors (CIDens.cpp:166-262)
    def get_one_rdm(self, spin, a, b):
        L = self.ham.L
        if a > b:
            a, b = b, a
        return float(self._one[spin * (L * (L + 1) // 2) + a * L + b - a * (a + 1) // 2])

    def get_two_rdm(self, spin, i, j, k, l):
        L = self.ham.L
        C2 = L * L
        tsz = C2 * (C2 + 1) // 2
        idx = lambda r, c: C2 * min(r, c) + max(r, c) - min(r, c) * (min(r, c) + 1) // 2  # noqa: E731
        if spin == 1:
            return float(self._two[tsz + idx(i * L + j, k * L + l)])
        if i == j or k == l:
            return 0.0
        sign = 1.0
        if i > j:
            i, j, sign = j, i, -sign
        if k > l:
            k, l, sign = l, k, -sign
        P = i * L + j - (i + 1) * (i + 2) // 2
        Q = k * L + l - (k + 1) * (k + 2) // 2
        return sign * float(self._two[spin * tsz + idx(P, Q)])

    def get_dens_energy(self):
        """CIDens::get_dens_energy (CIDens.cpp:898-932) from the dense RDMs."""
        g, G = self.density_dense()
        h, V = self.ham.h, self.ham.V
        e1 = float(np.sum(h * (g[0] + g[1])))
        e2 = 0.5 * float(np.sum(V * (G[0] + 2 * G[1] + G[2])))
        return e1 + e2 + self.ham.econst

    @property
    def ptr(self):
        return self.space.ptr


class FCI(CIMethod):
    kind = CF_FCI

    def get_name(self):
        return "FCI"


class DOCI(CIMethod):
    kind = CF_DOCI

    def get_name(self):
        return "DOCI"


def read_integrals(path: str) -> dict:
    """Hamiltonian::read_file (Hamiltonian.cpp:707-885): the psi4 integral dump
    as dense h[L,L], physics V[L,L,L,L], econst, nalpha, nbeta and the orbital
    irreps."""
    L, na, nb, ec = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    _check(lib().cf_read_integrals(path.encode(), C.byref(L), C.byref(na), C.byref(nb), C.byref(ec), None, None,
                                   None))
    n = L.value
    h = np.zeros((n, n))
    V = np.zeros((n, n, n, n))
    ir = (C.c_int * n)()
    _check(lib().cf_read_integrals(path.encode(), C.byref(L), C.byref(na), C.byref(nb), C.byref(ec), _d(h), _d(V),
                                   ir))
    return {"L": n, "nalpha": na.value, "nbeta": nb.value, "econst": ec.value, "h": h, "V": V,
            "irreps": np.array(list(ir), dtype=int)}


def read_determinants(path: str):
    """Permutator_File::read_file (Permutator.cpp:179-219): the (up, down)
    strings of a determinant file, as uint64 arrays in file order."""
    cnt = C.c_uint64()
    _check(lib().cf_read_determinants(path.encode(), 0, None, None, C.byref(cnt)))
    up = np.zeros(cnt.value, dtype=np.uint64)
    down = np.zeros(cnt.value, dtype=np.uint64)
    _check(lib().cf_read_determinants(path.encode(), cnt.value, up.ctypes.data_as(u64p), down.ctypes.data_as(u64p),
                                      C.byref(cnt)))
    return up, down


class FCI_File(CIMethod):
    """Selected CI over a determinant list (CIMethod.h:164-174,
    CIMethod.cpp:890-917): FCI_File(ham, "dets.txt") or FCI_File(ham, (up, down)).
    Rows follow the list order; construct_density gives DensFILE's RDMs
    (CIDens.cpp:1468-1526) for L <= 32 and distinct determinants."""

    kind = CF_DETS

    def __init__(self, ham: Hamiltonian, dets, value_prec: int = CF_F64, build: bool = True, zero: float = 1e-14):
        if isinstance(dets, (str, bytes, os.PathLike)):
            self.filename = os.fspath(dets)
            up, down = read_determinants(self.filename)
        else:
            self.filename = None
            up, down = dets
        self.up = np.ascontiguousarray(up, dtype=np.uint64)
        self.down = np.ascontiguousarray(down, dtype=np.uint64)
        super().__init__(ham, value_prec=value_prec, build=build, zero=zero)

    def _create_space(self):
        _check(lib().cf_space_create_dets(self.ham.ptr, len(self.up), self.up.ctypes.data_as(u64p),
                                          self.down.ctypes.data_as(u64p), C.byref(self.space.ptr)))

    def get_name(self):
        return "FCI_File"


class FCI_Direct(CIMethod):
    """FCI with the on-the-fly sigma instead of a stored matrix -- the role of
    the reference's CI_Big (CIMethod.h:176-189, CIMethod.cpp:943-1116), for
    spaces too large to store."""

    kind = CF_FCI

    def __init__(self, ham: Hamiltonian, value_prec: int = CF_F64):
        super().__init__(ham, value_prec=value_prec, build=False)
        self.sig = _Handle("cf_sigma_destroy", self.space)
        _check(lib().cf_sigma_create(self.ptr, value_prec, C.byref(self.sig.ptr)))

    def get_name(self):
        return "FCI_Direct"

    def set_ham(self, ham: Hamiltonian, build: bool = True):
        """CIMethod::set_ham: the space stays resident, the sigma's integral
        tables are made again for the new Hamiltonian."""
        self.sig = None
        super().set_ham(ham, build=False)
        self.sig = _Handle("cf_sigma_destroy", self.space)
        _check(lib().cf_sigma_create(self.ptr, self._value_prec, C.byref(self.sig.ptr)))

    def _new_space(self, ham: Hamiltonian, build: bool):
        FCI_Direct.__init__(self, ham, value_prec=self._value_prec)

    def set_threads(self, threads: int):
        """Workgroup size of the alpha-beta sigma kernel (128, 256, 512, 768, 1024)."""
        _check(lib().cf_sigma_set_threads(self.sig.ptr, threads))

    def set_kernel(self, kernel: int):
        """Alpha-beta sigma kernel: 1 staged slices, 2 prefetched slice tables."""
        _check(lib().cf_sigma_set_kernel(self.sig.ptr, kernel))

    def build_ci(self):
        raise CIFlowError(CF_ERR_STATE, "FCI_Direct stores no matrix")

    def mvprod(self, x, precision: int = CF_F64):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().cf_sigma_apply(self.sig.ptr, precision, _d(x), _d(y)))
        return y

    def solve(self, precision: int = CF_F64, want_vec: bool = True, **kw):
        opts = solve_opts(precision=precision, **kw)
        info = SolveInfo()
        vec = np.zeros(self.n) if want_vec else None
        _check(lib().cf_solve_direct(self.sig.ptr, C.byref(opts), _d(vec) if want_vec else None, C.byref(info)))
        self._info, self._eigvec, self._eigval = info, vec, info.eigval
        self._one = self._two = None
        return info

    def construct_density(self, trdm: bool = False):
        if trdm:
            raise CIFlowError(CF_ERR_UNSUPPORTED, "FCI_Direct computes the 1-RDM only")
        L = self.ham.L
        one = np.zeros(L * (L + 1))
        c = self._eigvec
        if c is None:
            c = np.zeros(self.n)
            _check(lib().cf_space_eigvec(self.ptr, _d(c)))
        _check(lib().cf_rdm1_direct(self.ptr, _d(np.ascontiguousarray(c)), _d(one)))
        self._one = one
        return one, None


class CI_Big(CIMethod):
    """Selected CI over a determinant list without storing its matrix
    (CIMethod.h:176-189, CIMethod.cpp:591-609, 943-1116): CI_Big(ham,
    "dets.txt") or CI_Big(ham, (up, down)).  ``stored_rows`` leading rows and
    columns are stored as FCI_File stores its matrix (the reference keeps as
    many as its memory budget allows); every other element is evaluated on
    the GPU in each product from the sorted determinant keys.  The solve is
    Lanczos (ARPACK semantics) on that product; construct_density gives
    DensFILE's RDMs, as for FCI_File."""

    kind = CF_DETS

    def __init__(self, ham: Hamiltonian, dets, stored_rows: int = 0):
        if isinstance(dets, (str, bytes, os.PathLike)):
            self.filename = os.fspath(dets)
            up, down = read_determinants(self.filename)
        else:
            self.filename = None
            up, down = dets
        self.up = np.ascontiguousarray(up, dtype=np.uint64)
        self.down = np.ascontiguousarray(down, dtype=np.uint64)
        self.stored_rows = int(stored_rows)
        super().__init__(ham, build=False)
        self.sig = _Handle("cf_sigma_destroy", self.space)
        _check(lib().cf_dets_sigma_create(self.ptr, self.stored_rows, C.byref(self.sig.ptr)))

    def _create_space(self):
        _check(lib().cf_space_create_dets(self.ham.ptr, len(self.up), self.up.ctypes.data_as(u64p),
                                          self.down.ctypes.data_as(u64p), C.byref(self.space.ptr)))

    def _new_space(self, ham: Hamiltonian, build: bool):
        CI_Big.__init__(self, ham, (self.up, self.down), self.stored_rows)

    def set_ham(self, ham: Hamiltonian, build: bool = True):
        """CIMethod::set_ham: a new list space and product for the new integrals."""
        self.sig = None
        self._new_space(ham, build)

    def get_name(self):
        return "CI_Big"

    def build_ci(self):
        raise CIFlowError(CF_ERR_STATE, "CI_Big stores at most its leading block, inside its product")

    def mvprod(self, x, precision: int = CF_F64):
        """CI_Big::mvprod (CIMethod.cpp:943-984)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        _check(lib().cf_sigma_apply(self.sig.ptr, precision, _d(x), _d(y)))
        return y

    def solve(self, precision: int = CF_F64, want_vec: bool = True, **kw):
        """CI_Big::solve / arpackDiagonalize (CIMethod.cpp:1000-1116)."""
        opts = solve_opts(precision=precision, **kw)
        info = SolveInfo()
        vec = np.zeros(self.n) if want_vec else None
        _check(lib().cf_solve_direct(self.sig.ptr, C.byref(opts), _d(vec) if want_vec else None, C.byref(info)))
        self._info, self._eigvec, self._eigval = info, vec, info.eigval
        self._one = self._two = None
        return info


class SolveServer(_Handle):
    """Continuously batched stream of rotated-integral ground-state solves
    (the set_ham -> build_ci -> solve loop of OrbitalOptimization::update_cim,
    OrbitalOptimization.cpp, driven by SimulatedAnnealing.cpp:83-87 and
    LocalMinimizer.cpp:73-235), served by ``slots`` worker streams on one GPU
    with HBM-footprint admission."""

    def __init__(self, slots: int = 4, hbm_budget: float = 0.0, device: int = 0):
        super().__init__("cf_server_destroy")
        _check(lib().cf_server_create(device, slots, float(hbm_budget), C.byref(self.ptr)))

    def close(self):
        """Stop the workers (each finishes the request it is running)."""
        if self.ptr:
            _check(lib().cf_server_destroy(self.ptr))
            self.ptr = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def _request(tag, kind, h, V, nup, ndown, U=None, econst=0.0, precision=CF_F64, ncv=0, tol=0.0, check_every=0,
                 tol_energy=0.0, direct=False, method=0):
        h = np.ascontiguousarray(h, dtype=np.float64)
        V = np.ascontiguousarray(V, dtype=np.float64)
        U = None if U is None else np.ascontiguousarray(U, dtype=np.float64)
        L = h.shape[0]
        assert h.shape == (L, L) and V.shape == (L, L, L, L) and (U is None or U.shape == (L, L))
        r = Request(tag, kind, L, nup, ndown, _d(h), _d(V), _d(U) if U is not None else None, econst, precision,
                    ncv, tol, check_every, tol_energy, int(bool(direct)), int(method))
        return r, (h, V, U)

    def submit(self, tag: int, kind: int, h, V, nup: int, ndown: int, U=None, econst: float = 0.0,
               precision: int = CF_F64, ncv: int = 0, tol: float = 0.0, check_every: int = 0,
               tol_energy: float = 0.0, direct: bool = False, method: int = 0):
        r, keep = self._request(tag, kind, h, V, nup, ndown, U, econst, precision, ncv, tol, check_every,
                                tol_energy, direct, method)
        _check(lib().cf_server_submit(self.ptr, C.byref(r)))  # arrays are copied inside

    @staticmethod
    def footprint(kind: int, L: int, nup: int, ndown: int, precision: int = CF_F64, ncv: int = 0,
                  direct: bool = False) -> int:
        z = np.zeros(1)
        r = Request(0, kind, L, nup, ndown, _d(z), _d(z), None, 0.0, precision, ncv, 0.0, 0, 0.0, int(bool(direct)), 0)
        b = C.c_uint64()
        _check(lib().cf_request_footprint(C.byref(r), C.byref(b)))
        return b.value

    def poll(self, max_results: int = 256) -> list[dict]:
        buf = (Result * max_results)()
        cnt = C.c_int()
        _check(lib().cf_server_poll(self.ptr, buf, max_results, C.byref(cnt)))
        return [buf[i].as_dict() for i in range(cnt.value)]

    def drain(self, timeout: float = 0.0):
        _check(lib().cf_server_drain(self.ptr, timeout))

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        _check(lib().cf_server_stats(self.ptr, *[C.byref(x) for x in v]))
        return dict(zip(["submitted", "finished", "hbm_in_use", "hbm_budget"], [x.value for x in v]))


def solve_doci_batch(requests, ctx: "Context | None" = None) -> list[dict]:
    """cf_solve_doci_batch: many small DOCI solves in one launch.  requests:
    dicts with h, V, npair and optionally U, econst, tol, tol_energy,
    check_every, tag; returns one result dict per request (SolveServer.poll's
    fields)."""
    ctx = ctx or default_context()
    reqs = (Request * len(requests))()
    keep = []
    for i, q in enumerate(requests):
        r, arrs = SolveServer._request(q.get("tag", i), CF_DOCI, q["h"], q["V"], q["npair"], q["npair"],
                                       q.get("U"), q.get("econst", 0.0), CF_F64, 0, q.get("tol", 0.0),
                                       q.get("check_every", 0), q.get("tol_energy", 0.0), False, CF_SOLVER_AUTO)
        reqs[i] = r
        keep.append(arrs)
    res = (Result * len(requests))()
    _check(lib().cf_solve_doci_batch(ctx.ptr, len(requests), reqs, res))
    return [res[i].as_dict() for i in range(len(requests))]


def solve_batch(requests, ctx: "Context | None" = None) -> list[dict]:
    """cf_solve_batch: batched Davidson solves, requests of one shape advanced
    in lock step.  requests: dicts with kind (CF_FCI / CF_DOCI), h, V, nup,
    ndown (or npair) and optionally U, econst, precision (CF_F32), ncv, tol,
    tol_energy, tag; FCI requests run on the on-the-fly sigma.  Returns one
    result dict per request (SolveServer.poll's fields)."""
    ctx = ctx or default_context()
    reqs = (Request * len(requests))()
    keep = []
    for i, q in enumerate(requests):
        nup = q.get("nup", q.get("npair"))
        ndown = q.get("ndown", q.get("npair"))
        r, arrs = SolveServer._request(q.get("tag", i), q.get("kind", CF_FCI), q["h"], q["V"], nup, ndown,
                                       q.get("U"), q.get("econst", 0.0), q.get("precision", CF_F32), q.get("ncv", 0),
                                       q.get("tol", 0.0), 0, q.get("tol_energy", 0.0), True,
                                       q.get("method", CF_SOLVER_AUTO))
        reqs[i] = r
        keep.append(arrs)
    res = (Result * len(requests))()
    _check(lib().cf_solve_batch(ctx.ptr, len(requests), reqs, res))
    return [res[i].as_dict() for i in range(len(requests))]


class RCILanczos:
    """dsaupd_-shaped reverse communication (host vectors): call step() until
    it returns 99; on 1 compute workd[n:2n] = A @ workd[0:n]."""

    def __init__(self, n: int, ncv: int = 42, tol: float = 1e-12, maxrestart: int = 0, seed: int = 20240601):
        self.n = n
        self.handle = _Handle("cf_rci_destroy")
        self.ptr = self.handle.ptr
        _check(lib().cf_rci_create(n, ncv, tol, maxrestart, seed, C.byref(self.handle.ptr)))
        self.workd = np.zeros(3 * n)
        self.ido = C.c_int(0)
        self.ipntr = (C.c_int * 11)()

    def step(self) -> int:
        _check(lib().cf_rci_step(self.ptr, C.byref(self.ido), _d(self.workd), self.ipntr))
        return self.ido.value

    def result(self):
        ev = C.c_double()
        vec = np.zeros(self.n)
        info = C.c_int()
        _check(lib().cf_rci_result(self.ptr, C.byref(ev), _d(vec), C.byref(info)))
        return ev.value, vec, info.value


def arpack_like(matvec, n: int, ncv: int = 42, tol: float = 1e-12):
    """The reference's arpackDiagonalize loop (SparseMatrix_CRS.cpp:441-449)
    over an arbitrary mvprod: returns (eigval, eigvec, info)."""
    r = RCILanczos(n, ncv, tol)
    while r.step() != 99:
        i0 = r.ipntr[0] - 1
        i1 = r.ipntr[1] - 1
        r.workd[i1:i1 + n] = matvec(r.workd[i0:i0 + n])
    return r.result()


from . import h5  # noqa: E402  (HDF5 files; its library loads on first use)
