"""The no-allocation entry point (appt_spgemm_compute_noalloc) against
appt_spgemm_compute through the C ABI: two calls in a row, the second on a
smaller input, outputs identical to the allocating entry point's and to the
oracle within the per-entry tolerance; the engine's device pool does not grow;
oversized inputs and short output buffers are refused cleanly."""
import ctypes as C
import sys

import numpy as np
import pytest

import appt_spgemm_amd as ap
from appt_spgemm_amd._abi import OPS, CsrView, lib
from tests._util import assert_product, ref

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DevBuf:
    """Caller-owned device buffer (plain hipMalloc through ctypes)."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(int(nbytes), 8)) == 0
        self.nbytes = nbytes

    def get(self, dtype, count):
        out = np.empty(max(count, 1), dtype)
        if count:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.itemsize * count, 2) == 0  # device -> host
        return out[:count]

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def noalloc(eng, A, op, cap):
    n = A.n
    rp, col, val = DevBuf(4 * (n + 1)), DevBuf(4 * cap), DevBuf(8 * cap)
    nnz = C.c_int64(-1)
    st = lib().appt_spgemm_compute_noalloc(eng.handle, C.byref(A.view), OPS[op], rp.ptr, col.ptr, val.ptr, cap,
                                           C.byref(nnz))
    return st, nnz.value, (rp, col, val)


def test_noalloc_matches_compute():
    eng = ap.Engine(0)
    try:
        assert eng.device_bytes() == 0
        cases = [("stencil27 g=24", ap.stencil27(24, 0.1, 5), "AA"),          # first call: larger input
                 ("er 3000", ap.erdos_renyi(3000, 6.0, 1), "AAt"),            # second call: smaller
                 ("rmat 11", ap.rmat(11, 16, seed=2), "AAtS"),
                 ("laplace 16", ap.laplace2d(16), "AA")]
        for name, m, op in cases:
            say("case noalloc", name, op, "n", m.n, "nnz", m.nnz)
            A = eng.upload(m)
            base = eng.device_bytes()
            try:
                want = ref(m, op)
                st, nnz, bufs = noalloc(eng, A, op, int(want[0][-1]))
                try:
                    assert st == 0, lib().appt_last_error()
                    assert eng.device_bytes() == base        # nothing allocated by the call
                    assert nnz == int(want[0][-1])
                    rp, col, val = (b.get(t, c) for b, t, c in zip(bufs, (np.int32, np.int32, np.float64),
                                                                    (m.n + 1, nnz, nnz)))
                    got = ap.HostMatrix(m.n, rp, col, val)
                    assert_product(got, want, exact=False)
                    R = eng.compute(A, op)                   # the allocating entry point
                    try:
                        h = R.download()
                    finally:
                        R.release()
                    np.testing.assert_array_equal(np.asarray(h.rowptr, np.int64), rp.astype(np.int64))
                    np.testing.assert_array_equal(h.col, col)
                    assert_product(ap.HostMatrix(m.n, rp, col, val), (h.rowptr, h.col, h.val, want[3]),
                                   exact=name != "rmat 11")
                finally:
                    for b in bufs:
                        b.free()
            finally:
                A.free()
    finally:
        eng.close()


def test_noalloc_refuses_cleanly():
    eng = ap.Engine(0)
    try:
        m = ap.stencil27(10, 0.0, 1)
        A = eng.upload(m)
        try:
            need = (5 * 10 - 6) ** 3
            st, nnz, bufs = noalloc(eng, A, "AA", need - 1)         # one entry short
            for b in bufs:
                b.free()
            assert ap._abi.STATUS_NAMES[st] == "APPT_ERR_OVERFLOW" and nnz == need
            st, nnz, bufs = noalloc(eng, A, "AA", need)             # the same call then fits
            for b in bufs:
                b.free()
            assert st == 0 and nnz == need
        finally:
            A.free()
        # 400k rows exceed the static scratch: refused before anything is read
        dummy = DevBuf(8)
        try:
            big = CsrView(400_000, 0, dummy.ptr, None, None)
            nnz = C.c_int64(0)
            st = lib().appt_spgemm_compute_noalloc(eng.handle, C.byref(big), OPS["AA"], dummy.ptr, None, None, 0,
                                                   C.byref(nnz))
        finally:
            dummy.free()
        assert ap._abi.STATUS_NAMES[st] == "APPT_ERR_UNSUPPORTED"
        assert b"static scratch" in lib().appt_last_error() or b"no-allocation" in lib().appt_last_error()
    finally:
        eng.close()
