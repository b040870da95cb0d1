"""CPU oracle (TEST INFRASTRUCTURE ONLY): ctypes wrapper of oracle/oracle.c.

Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
package, and only as the checker or the timed CPU baseline.  The product
library never calls it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
OPS = {"AA": 0, "AAt": 1, "AAtS": 2}
_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE), "liboracle.so"], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        i32p, i64p, f64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)
        L.oracle_spgemm.restype = C.c_int
        L.oracle_spgemm.argtypes = [C.c_int32, i32p, i32p, f64p, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(i64p), C.POINTER(i32p), C.POINTER(f64p),
                                    C.POINTER(f64p)]
        L.oracle_gustavson_range.restype = C.c_int
        L.oracle_gustavson_range.argtypes = [C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p, C.c_int,
                                             C.c_int32, C.c_int32, i64p, i32p, f64p, f64p, C.c_int]
        L.oracle_transpose.restype = C.c_int
        L.oracle_transpose.argtypes = [C.c_int32, i32p, i32p, f64p, i32p, i32p, f64p]
        L.oracle_row_products.restype = None
        L.oracle_row_products.argtypes = [C.c_int32, i32p, i32p, i32p, i64p]
        L.oracle_caps.restype = None
        L.oracle_caps.argtypes = [C.c_int32, i64p, C.c_int32, i32p, i64p]
        L.oracle_free.restype = None
        L.oracle_free.argtypes = [C.c_void_p]
        L.oracle_checksum.restype = None
        L.oracle_checksum.argtypes = [f64p, C.c_int64, f64p]
        L.oracle_descriptors.restype = C.c_int
        L.oracle_descriptors.argtypes = [C.c_int32, i32p, i32p, f64p, f64p]
        L.oracle_pick_ours.restype = C.c_int
        L.oracle_pick_ours.argtypes = [C.c_double, C.c_double, C.c_int64]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _arr(rowptr, col, val):
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    cl = np.ascontiguousarray(col, dtype=np.int32)
    vl = np.ascontiguousarray(val, dtype=np.float64)
    return rp, cl, vl


def spgemm(n, rowptr, col, val, op="AA", threads=0, want_scale=True):
    """C = op(A) with structural semantics.  Returns (rowptr int64, col int32,
    val float64, scale float64 or None)."""
    rp, cl, vl = _arr(rowptr, col, val)
    L = lib()
    cp, cc, cv, cs = C.POINTER(C.c_int64)(), C.POINTER(C.c_int32)(), C.POINTER(C.c_double)(), \
        C.POINTER(C.c_double)()
    rc = L.oracle_spgemm(n, _p(rp, C.c_int32), _p(cl, C.c_int32), _p(vl, C.c_double), OPS[op],
                         threads, 1 if want_scale else 0, C.byref(cp), C.byref(cc), C.byref(cv),
                         C.byref(cs))
    if rc != 0:
        raise MemoryError("oracle_spgemm failed")
    rows = np.ctypeslib.as_array(cp, shape=(n + 1,)).copy()
    m = int(rows[-1])
    cols = np.ctypeslib.as_array(cc, shape=(max(m, 1),))[:m].copy()
    vals = np.ctypeslib.as_array(cv, shape=(max(m, 1),))[:m].copy()
    scale = np.ctypeslib.as_array(cs, shape=(max(m, 1),))[:m].copy() if want_scale else None
    for ptr in (cp, cc, cv) + ((cs,) if want_scale else ()):
        L.oracle_free(C.cast(ptr, C.c_void_p))
    return rows, cols, vals, scale


def transpose(n, rowptr, col, val):
    rp, cl, vl = _arr(rowptr, col, val)
    tp = np.empty(n + 1, np.int32)
    tc = np.empty(max(len(cl), 1), np.int32)
    tv = np.empty(max(len(cl), 1), np.float64)
    lib().oracle_transpose(n, _p(rp, C.c_int32), _p(cl, C.c_int32), _p(vl, C.c_double),
                           _p(tp, C.c_int32), _p(tc, C.c_int32), _p(tv, C.c_double))
    return tp, tc[:len(cl)], tv[:len(cl)]


def rows_range(n, rowptr, col, val, r0, r1, op="AA", threads=0, b=None, want_scale=False):
    """Rows [r0, r1) of op(A) for op in {AA, AAt} (the CPU-baseline sample).
    ``b`` may pass a prebuilt (rowptr, col, val) of B.  Returns (rowptr, col,
    val) or, with want_scale, (rowptr, col, val, sum |a*b| per entry)."""
    rp, cl, vl = _arr(rowptr, col, val)
    if b is None:
        b = (rp, cl, vl) if op == "AA" else transpose(n, rp, cl, vl)
    bp, bc, bv = _arr(*b)
    L = lib()
    cnt = np.empty(r1 - r0, np.int64)
    args = [n, _p(rp, C.c_int32), _p(cl, C.c_int32), _p(vl, C.c_double), _p(bp, C.c_int32),
            _p(bc, C.c_int32), _p(bv, C.c_double), 0, r0, r1]
    if L.oracle_gustavson_range(*args, _p(cnt, C.c_int64), None, None, None, threads):
        raise MemoryError("oracle_gustavson_range failed")
    off = np.zeros(r1 - r0 + 1, np.int64)
    np.cumsum(cnt, out=off[1:])
    m = int(off[-1])
    cols = np.empty(max(m, 1), np.int32)
    vals = np.empty(max(m, 1), np.float64)
    start = off[:-1].copy()
    scale = np.empty(max(m, 1), np.float64) if want_scale else None
    if L.oracle_gustavson_range(*args, _p(start, C.c_int64), _p(cols, C.c_int32),
                                _p(vals, C.c_double), _p(scale, C.c_double) if want_scale else None, threads):
        raise MemoryError("oracle_gustavson_range failed")
    if want_scale:
        return off, cols[:m], vals[:m], scale[:m]
    return off, cols[:m], vals[:m]


def row_products(n, rowptr, col, b_rowptr):
    rp, cl = np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(col, np.int32)
    bp = np.ascontiguousarray(b_rowptr, np.int32)
    f = np.empty(n, np.int64)
    lib().oracle_row_products(n, _p(rp, C.c_int32), _p(cl, C.c_int32), _p(bp, C.c_int32),
                              _p(f, C.c_int64))
    return f


def caps(f, sh_cap=1024):
    f = np.ascontiguousarray(f, np.int64)
    rc = np.empty(len(f), np.int32)
    gc = np.empty(len(f), np.int64)
    lib().oracle_caps(len(f), _p(f, C.c_int64), sh_cap, _p(rc, C.c_int32), _p(gc, C.c_int64))
    return rc, gc


def checksum(vals):
    v = np.ascontiguousarray(vals, np.float64)
    out = np.empty(3, np.float64)
    lib().oracle_checksum(_p(v, C.c_double), len(v), _p(out, C.c_double))
    return tuple(float(x) for x in out)


def descriptors(n, rowptr, col, val):
    rp, cl = np.ascontiguousarray(rowptr, np.int32), np.ascontiguousarray(col, np.int32)
    vl = np.ascontiguousarray(val, np.float64)
    out = np.empty(9, np.float64)
    if lib().oracle_descriptors(n, _p(rp, C.c_int32), _p(cl, C.c_int32), _p(vl, C.c_double), _p(out, C.c_double)):
        raise MemoryError("oracle_descriptors failed")
    keys = ["deg_mean", "deg_cv", "deg_max", "fill_aa", "fill_aat", "psym", "products_aa",
            "products_aat", "bandwidth"]
    return dict(zip(keys, out.tolist()))


def pick_ours(dcv, fill, nnz):
    return bool(lib().oracle_pick_ours(dcv, fill, nnz))
