"""GPU test of cf_rotate_integrals_noalloc, the integral rotation that
allocates nothing (its work array is a per-GPU device array in the library).

Two calls in a row on one context -- FCI-sized L = 12 with 4 Hamiltonians, then
a smaller L = 8 with 2 -- must give exactly the output of the allocating
cf_rotate_integrals on the same inputs (same kernels, same order: bitwise);
a batch beyond the work array is rejected with CF_ERR_ARG before any launch.
Device buffers for the inputs and outputs come from the HIP runtime the
library itself is linked against (test plumbing only).
"""
import ctypes as C
import sys

import numpy as np
import pytest

import ciflow_amd as CA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libciflow_amd.so loaded."""

    def __init__(self):
        CA.lib()
        path = next(ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64" in ln)
        self.rt = C.CDLL(path)
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        assert self.rt.hipMemset(p, 0xFF, nbytes) == 0   # NaN pattern: outputs must be written
        return p

    def to_dev(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def to_host(self, p, shape):
        out = np.empty(shape)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def inputs(L, batch, seed0):
    h = np.empty((batch, L, L))
    V = np.empty((batch, L, L, L, L))
    U = np.empty((batch, L, L))
    for b in range(batch):
        h[b], V[b] = CA.gen_integrals(L, seed0 + b)
        U[b] = CA.expm_skew(CA.gen_rotation(L, seed0 + b, 0.05))
    return h, V, U


def test_noalloc_rotation_equals_allocating_rotation_on_two_calls():
    ctx = CA.Context(0)
    hip = Hip()
    try:
        for L, batch, seed0 in [(12, 4, 100), (8, 2, 200)]:   # the second call is smaller
            say(f"noalloc rotation L={L} batch={batch}")
            h, V, U = inputs(L, batch, seed0)
            ref_h, ref_V = CA.rotate_integrals(h, V, U, ctx)
            dh, dV, dU = hip.to_dev(h), hip.to_dev(V), hip.to_dev(U)
            oh, oV = hip.alloc(h.nbytes), hip.alloc(V.nbytes)
            CA._check(CA.lib().cf_rotate_integrals_noalloc(ctx.ptr, L, batch, dh, dV, dU, oh, oV))
            ctx.synchronize()
            got_h, got_V = hip.to_host(oh, h.shape), hip.to_host(oV, V.shape)
            assert np.array_equal(got_h, ref_h)
            assert np.array_equal(got_V, ref_V)
            hip.free()
    finally:
        hip.free()


def test_noalloc_rotation_rejects_batch_beyond_work_array():
    ctx = CA.Context(0)
    one = C.c_void_p(1)   # never dereferenced: the size check comes first
    rc = CA.lib().cf_rotate_integrals_noalloc(ctx.ptr, 24, 7, one, one, one, one, one)   # 7 * 24^4 > 2^21
    assert rc == CA.CF_ERR_ARG
    assert b"exceeds" in CA.lib().cf_last_error()
