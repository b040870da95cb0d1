"""A*A^T as a triangle-only Gram (op AAtS): only the products of one
triangle are formed (each A entry's column suffix from its own row, or the
prefix up to it), then the off-diagonal entries are mirrored.  Every result
must equal the oracle's AAtS (bit-exact where rows accumulate in A-row order),
be bit-symmetric, and agree with the full-rows path (tuning gram = -1) and
with A*A^T."""
import sys

import numpy as np
import pytest

import appt_spgemm_amd as ap
import oracle
from tests._util import TOL, assert_product, messy, ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def engines():
    up = ap.Engine(0, tuning={"gram": 1})
    low = ap.Engine(0, tuning={"gram": 2})
    f = ap.Engine(0, tuning={"gram": -1})
    low64 = ap.Engine(0, tuning={"gram": 2, "rowptr64_min_nnz": 1})
    yield up, low, f, low64
    for e in (up, low, f, low64):
        e.close()


def host(n, rp, col, val):
    return ap.HostMatrix(n, np.asarray(rp, np.int32), np.asarray(col, np.int32), np.asarray(val, np.float64))


def cases():
    yield "n1", host(1, [0, 1], [0], [3.0]), True
    yield "all_empty", host(10, [0] * 11, [], []), True
    yield "empty_rows", host(6, [0, 2, 2, 3, 3, 5, 5], [1, 4, 0, 2, 5], [1.0, 2.0, 3.0, 4.0, 5.0]), True
    yield "unsorted", host(4, [0, 3, 4, 6, 7], [3, 0, 2, 1, 3, 0, 2], [1, 2, 3, 4, 5, 6, 7.0]), True
    yield "duplicates", host(3, [0, 3, 4, 6], [1, 1, 0, 2, 0, 0], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]), False
    yield "last_col_only", host(5, [0, 1, 2, 3, 4, 5], [4, 4, 4, 4, 4], [1.0, 2.0, 3.0, 4.0, 5.0]), True
    yield "laplace2d_40", ap.laplace2d(40), True
    yield "stencil27_12_drop", ap.stencil27(12, 0.1, 3), True
    yield "banded_5000", ap.banded(5000, 12, 0.7, 4), True
    yield "er_20000", ap.erdos_renyi(20000, 6.0, 5), True
    yield "messy_3000", messy(3000, 6), False
    yield "rmat_11", ap.rmat(11, 16, seed=7), False
    yield "rmat_13", ap.rmat(13, 16, seed=8), False


def sym_exact(got):
    C = ap.HostMatrix(got.n, got.rowptr, got.col, got.val).to_scipy()
    D = (C - C.T).tocoo()
    assert np.all(D.data == 0.0), "C is not bit-symmetric"
    T = C.T.tocsr()
    T.sort_indices()
    np.testing.assert_array_equal(T.indptr, np.asarray(got.rowptr, np.int64))
    np.testing.assert_array_equal(T.indices, got.col)


@pytest.mark.parametrize("name,m,exact", list(cases()), ids=[c[0] for c in cases()])
def test_gram_matches_oracle(engines, name, m, exact):
    up, low, f, low64 = engines
    r = ref(m, "AAtS")
    full = f.multiply(m, "AAtS")
    for tri, g in (("upper", up), ("lower", low)):
        say("case gram", tri, name, m.n, m.nnz)
        got = g.multiply(m, "AAtS")
        assert_product(got, r, exact=exact)
        sym_exact(got)
        np.testing.assert_array_equal(full.rowptr, got.rowptr)
        np.testing.assert_array_equal(full.col, got.col)
        assert np.all(np.abs(full.val - got.val) <= TOL * r[3])
    # int64 row pointers and the int64 copy of the expanded offsets
    got64 = low64.multiply(m, "AAtS")
    assert_product(got64, r, exact=exact)


@pytest.mark.parametrize("tri", [0, 1])
def test_gram_equals_aat_and_counts_full_products(engines, tri):
    g = engines[tri]
    m = ap.stencil27(20, 0.1, 9)
    A = g.upload(m)
    try:
        say("case gram vs AAt, products")
        g.set_timing(True)
        R = g.compute(A, "AAtS")
        t = g.phase_times()
        got = R.download()
        R.release()
        S = g.compute(A, "AAt")
        aat = S.download()
        S.release()
        np.testing.assert_array_equal(got.rowptr, aat.rowptr)
        np.testing.assert_array_equal(got.col, aat.col)
        # rows accumulate in A-row order on both paths: c_ij = c_ji bit for bit
        np.testing.assert_array_equal(got.val, aat.val)
        d = oracle.descriptors(m.n, m.rowptr, m.col, m.val)
        assert t["products"] == int(d["products_aat"])
        assert t["nnz_c"] == got.nnz
    finally:
        A.free()
        g.set_timing(False)


def test_gram_default_and_panels(engines):
    # the default engine picks its path; a panel of AAtS holds complete rows
    eng = ap.Engine(0)
    m = ap.rmat(12, 16, seed=4)
    A = eng.upload(m)
    try:
        say("case gram default + panels")
        R = eng.compute(A, "AAtS")
        whole = R.download()
        R.release()
        assert_product(whole, ref(m, "AAtS"), exact=False)
        rp = eng.symbolic_rowptr(A, "AAtS")
        np.testing.assert_array_equal(rp, np.asarray(whole.rowptr, np.int64))
        for r0, r1 in ((0, 700), (700, 2100), (2100, m.n)):
            P = eng.compute_rows(A, "AAtS", r0, r1)
            part = P.download()
            P.release()
            lo, hi = whole.rowptr[r0], whole.rowptr[r1]
            np.testing.assert_array_equal(part.rowptr, whole.rowptr[r0:r1 + 1] - lo)
            np.testing.assert_array_equal(part.col, whole.col[lo:hi])
            scale = ref(m, "AAtS")[3][lo:hi]
            assert np.all(np.abs(part.val - whole.val[lo:hi]) <= TOL * scale)
    finally:
        A.free()
        eng.close()
