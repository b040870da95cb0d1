"""Shared parity helpers for the tests."""
import numpy as np

import oracle

TOL = 1e-12  # per-entry: |c - c_ref| <= TOL * sum_k |a_ik * b_kj|


def to_csr(n, rp, col, val):
    return (np.asarray(rp, np.int64), np.asarray(col, np.int32), np.asarray(val, np.float64))


def ref(m, op):
    """Oracle product of a HostMatrix-like object (n, rowptr, col, val)."""
    return oracle.spgemm(m.n, m.rowptr, m.col, m.val, op)


def assert_structure(got_rp, got_col, ref_rp, ref_col):
    got_rp = np.asarray(got_rp, np.int64)
    np.testing.assert_array_equal(got_rp, np.asarray(ref_rp, np.int64), err_msg="row pointers differ")
    np.testing.assert_array_equal(np.asarray(got_col), np.asarray(ref_col), err_msg="column indices differ")


def assert_values(got, ref_val, scale, exact=False, tol=TOL):
    got = np.asarray(got, np.float64)
    ref_val = np.asarray(ref_val, np.float64)
    assert got.shape == ref_val.shape
    if exact:
        bad = np.nonzero(got != ref_val)[0]
        assert bad.size == 0, f"{bad.size} values not bit-equal, first at {bad[:5]}: {got[bad[:5]]} vs {ref_val[bad[:5]]}"
    err = np.abs(got - ref_val)
    lim = tol * np.asarray(scale, np.float64)
    bad = np.nonzero(err > lim)[0]
    assert bad.size == 0, f"{bad.size} values outside {tol}*sum|ab|, first {bad[:5]}: {err[bad[:5]]} > {lim[bad[:5]]}"


def assert_product(got, ref_tuple, exact=False):
    rp, col, val, scale = ref_tuple
    assert_structure(got.rowptr, got.col, rp, col)
    assert_values(got.val, val, scale, exact=exact)


def rows_sorted(rp, col):
    rp = np.asarray(rp, np.int64)
    d = np.diff(col.astype(np.int64))
    starts = np.zeros(len(col), bool)
    starts[rp[1:-1][rp[1:-1] < len(col)]] = True
    return bool(np.all((d > 0) | starts[1:]))


def messy(n: int, seed: int, mean_len: int = 6):
    """A square CSR matrix the way files and assemblers hand them over: rows
    in random column order, duplicate positions (half of them cancelling to
    zero), explicit zeros, and a large symmetric part (40 % of the entries
    mirrored).  Returns an engine HostMatrix."""
    import appt_spgemm_amd as ap
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 2 * mean_len + 1, n)
    r = np.repeat(np.arange(n), k)
    c = rng.integers(0, n, r.size)
    v = rng.uniform(-1, 1, r.size)
    mir = rng.random(r.size) < 0.4
    dup = rng.random(r.size) < 0.1
    cancel = rng.random(int(dup.sum())) < 0.5
    vd = np.where(cancel, -v[dup], rng.uniform(-1, 1, int(dup.sum())))
    nz = max(1, n // 5)
    zr, zc = rng.integers(0, n, nz), rng.integers(0, n, nz)
    R = np.concatenate([r, c[mir], r[dup], zr])
    Cc = np.concatenate([c, r[mir], c[dup], zc])
    V = np.concatenate([v, rng.uniform(-1, 1, int(mir.sum())), vd, np.zeros(nz)])
    order = np.lexsort((rng.random(R.size), R))     # by row, random within a row
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, R + 1, 1)
    return ap.HostMatrix(n, np.cumsum(rp).astype(np.int32), Cc[order].astype(np.int32), V[order].astype(np.float64))


def scipy_psym(m):
    """psym as src/fetch_and_prep.py:39-46 computes it: merge duplicates,
    keep nonzero values, |P & P^T| / |P|."""
    import scipy.sparse as sp
    A = sp.csr_matrix((m.val.copy(), m.col.copy(), m.rowptr.copy()), shape=(m.n, m.n))
    A.sum_duplicates()
    P = (A != 0).astype(np.int8)
    return P.multiply(P.T).nnz / max(1, P.nnz)
