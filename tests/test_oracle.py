"""Pins the CPU oracle against closed forms, dense brute force (golden
fixtures), scipy pattern products and algebraic invariants."""
from pathlib import Path

import numpy as np
import pytest

import appt_spgemm_amd as ap
import oracle
from tests._util import assert_structure, assert_values, rows_sorted

GOLD = Path(__file__).resolve().parent / "golden" / "golden_small.npz"


def golden_cases():
    z = np.load(GOLD)  # allow_pickle defaults to False
    for t in range(int(z["ncases"])):
        yield t, int(z[f"c{t}_n"]), z[f"c{t}_rowptr"], z[f"c{t}_col"], z[f"c{t}_val"], z


@pytest.mark.parametrize("op", ["AA", "AAt", "AAtS"])
def test_golden_bruteforce(op):
    for t, n, rp, col, val, z in golden_cases():
        key = "AAt" if op == "AAtS" else op
        crp, ccol, cval, scale = oracle.spgemm(n, rp, col, val, op)
        assert_structure(crp, ccol, z[f"c{t}_{key}_rowptr"], z[f"c{t}_{key}_col"])
        assert_values(cval, z[f"c{t}_{key}_val"], scale + 1e-300)


def c1_expected(g):
    """Closed form of A^2 for the 2D 5-point Laplacian (4 / -1): diagonal
    16 + deg(i); grid neighbours -8; distance 2 along an axis 1; diagonal
    neighbours 2 (two paths)."""
    n = g * g
    rows, cols, vals = [], [], []
    for y in range(g):
        for x in range(g):
            i = x + g * y
            deg = (x > 0) + (x < g - 1) + (y > 0) + (y < g - 1)
            for dy in range(-2, 3):
                for dx in range(-2, 3):
                    d = abs(dx) + abs(dy)
                    xx, yy = x + dx, y + dy
                    if d > 2 or not (0 <= xx < g and 0 <= yy < g):
                        continue
                    if d == 0:
                        v = 16.0 + deg
                    elif d == 1:
                        v = -8.0
                    elif dx == 0 or dy == 0:
                        v = 1.0
                    else:
                        v = 2.0
                    rows.append(i); cols.append(xx + g * yy); vals.append(v)
    order = np.lexsort((np.array(cols), np.array(rows)))
    r, c, v = np.array(rows)[order], np.array(cols)[order], np.array(vals)[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, r + 1, 1)
    return np.cumsum(rp), c, v


@pytest.mark.parametrize("op", ["AA", "AAt", "AAtS"])
def test_laplace_closed_form(op):
    m = ap.laplace2d(64)
    assert m.nnz == 20224
    rp, col, val, _ = oracle.spgemm(m.n, m.rowptr, m.col, m.val, op)
    assert rp[-1] == 51972
    erp, ecol, eval_ = c1_expected(64)
    assert_structure(rp, col, erp, ecol)
    np.testing.assert_array_equal(val, eval_)
    assert sorted(set(val.tolist())) == [-8.0, 1.0, 2.0, 18.0, 19.0, 20.0]


@pytest.mark.parametrize("g", [3, 5, 8])
def test_stencil_closed_form_counts(g):
    m = ap.stencil27(g, 0.0, 7)
    assert m.nnz == (3 * g - 2) ** 3
    f = oracle.row_products(m.n, m.rowptr, m.col, m.rowptr)
    assert int(f.sum()) == (9 * g - 10) ** 3
    for op in ("AA", "AAt", "AAtS"):
        rp, col, _, _ = oracle.spgemm(m.n, m.rowptr, m.col, m.val, op)
        assert int(rp[-1]) == (5 * g - 6) ** 3
        assert rows_sorted(rp, col)


def test_scipy_pattern_product():
    sp = pytest.importorskip("scipy.sparse")
    for m in (ap.erdos_renyi(3000, 6.0, 3), ap.rmat(11, 8, seed=5), ap.banded(2000, 9, 0.7, 2)):
        A = m.to_scipy()
        P = A.copy(); P.data[:] = 1.0
        for op in ("AA", "AAt"):
            B = P if op == "AA" else P.T.tocsr()
            patt = (P @ B).tocsr(); patt.sort_indices()
            rp, col, val, scale = oracle.spgemm(m.n, m.rowptr, m.col, m.val, op)
            assert_structure(rp, col, patt.indptr, patt.indices)
            num = (A @ (A if op == "AA" else A.T.tocsr())).tocsr()
            dense_at = np.asarray(num[np.repeat(np.arange(m.n), np.diff(rp)), col]).ravel()
            assert_values(val, dense_at, scale + 1e-300)


def test_symmetric_invariants():
    m = ap.laplace2d(20)
    aa = oracle.spgemm(m.n, m.rowptr, m.col, m.val, "AA")
    aat = oracle.spgemm(m.n, m.rowptr, m.col, m.val, "AAt")
    for a, b in zip(aa[:3], aat[:3]):
        np.testing.assert_array_equal(a, b)
    u = ap.stencil27(7, 0.1, 3)   # unsymmetric: AAtS == AAt up to association
    aat = oracle.spgemm(u.n, u.rowptr, u.col, u.val, "AAt")
    aats = oracle.spgemm(u.n, u.rowptr, u.col, u.val, "AAtS")
    assert_structure(aats[0], aats[1], aat[0], aat[1])
    assert_values(aats[2], aat[2], aat[3])
    # mirrored values are exactly symmetric
    C = ap.HostMatrix(u.n, aats[0], aats[1], aats[2]).to_scipy()
    assert (C != C.T).nnz == 0


def test_row_sum_invariant():
    for m in (ap.stencil27(6, 0.1, 11), ap.erdos_renyi(500, 5.0, 1)):
        A = m.to_scipy()
        ones = np.ones(m.n)
        for op, B in (("AA", A), ("AAt", A.T.tocsr())):
            rp, col, val, scale = oracle.spgemm(m.n, m.rowptr, m.col, m.val, op)
            C = ap.HostMatrix(m.n, rp, col, val).to_scipy()
            np.testing.assert_allclose(C @ ones, A @ (B @ ones), rtol=1e-12, atol=1e-12)


def test_cancellation_kept():
    # [[1,1],[1,-1]] * its transpose: off-diagonals cancel to 0 but stay stored
    rp = np.array([0, 2, 4], np.int32); col = np.array([0, 1, 0, 1], np.int32)
    val = np.array([1.0, 1.0, 1.0, -1.0])
    crp, ccol, cval, _ = oracle.spgemm(2, rp, col, val, "AAt")
    assert crp.tolist() == [0, 2, 4] and ccol.tolist() == [0, 1, 0, 1]
    assert cval.tolist() == [2.0, 0.0, 0.0, 2.0]


def test_value_order_is_storage_order():
    # the oracle adds contributions in A-row storage order: with the entries of
    # row 0 permuted, (1e16 + 1) - 1e16 = 0 but (1e16 - 1e16) + 1 = 1 in fp64
    rp = np.array([0, 3, 4, 5, 6], np.int32)
    val_b = np.array([1.0, 1.0, 1.0])
    for order, expect in (([1, 2, 3], 0.0), ([1, 3, 2], 1.0)):
        col = np.array(order + [0, 0, 0], np.int32)
        v = np.concatenate([np.array([{1: 1e16, 2: 1.0, 3: -1e16}[k] for k in order]), val_b])
        crp, ccol, cval, _ = oracle.spgemm(4, rp, col, v, "AA")
        assert ccol[:1].tolist() == [0]
        assert cval[0] == expect


def test_row_range_matches_full():
    m = ap.stencil27(6, 0.1, 5)
    for op in ("AA", "AAt"):
        rp, col, val, _ = oracle.spgemm(m.n, m.rowptr, m.col, m.val, op)
        off, c2, v2 = oracle.rows_range(m.n, m.rowptr, m.col, m.val, 17, 101, op)
        np.testing.assert_array_equal(off + rp[17], rp[17:102])
        np.testing.assert_array_equal(c2, col[rp[17]:rp[101]])
        np.testing.assert_array_equal(v2, val[rp[17]:rp[101]])


def test_caps_rule_matches_reference_sizing():
    # k_caps: want = np2((4f)/3 + 1) (min 2); shared iff want <= 1024 (f <= 767)
    rc, gc = oracle.caps(np.array([0, 1, 2, 767, 768, 5000], np.int64))
    assert rc.tolist() == [0, 2, 4, 1024, 0, 0]
    assert gc.tolist() == [0, 0, 0, 0, 2048, 8192]


def test_checksum_sequential():
    v = np.array([1.0, -2.0, 3.5])
    assert oracle.checksum(v) == (2.5, 6.5, 1 + 4 + 12.25)


def test_descriptors_against_numpy():
    sp = pytest.importorskip("scipy.sparse")
    for m in (ap.rmat(10, 8, seed=2), ap.laplace2d(9), ap.erdos_renyi(700, 4.0, 9)):
        d = oracle.descriptors(m.n, m.rowptr, m.col, m.val)
        deg = np.diff(m.rowptr).astype(np.float64)
        assert d["deg_mean"] == pytest.approx(deg.mean())
        assert d["deg_cv"] == pytest.approx(deg.std() / deg.mean())
        assert d["fill_aa"] == pytest.approx(deg[m.col].sum() / m.nnz)
        cc = np.bincount(m.col, minlength=m.n).astype(np.float64)
        assert d["fill_aat"] == pytest.approx((cc * cc).sum() / m.nnz)
        P = m.to_scipy(); P.data[:] = 1
        assert d["psym"] == pytest.approx(P.multiply(P.T).nnz / P.nnz)
        rows = np.repeat(np.arange(m.n), np.diff(m.rowptr))
        assert d["bandwidth"] == np.abs(rows - m.col).max()


def test_descriptors_psym_numeric_pattern():
    # src/fetch_and_prep.py:39-46 merges duplicates and drops zeros before the
    # symmetry count; stored row lengths still give the degree statistics
    from tests._util import messy, scipy_psym
    for seed in range(4):
        m = messy(300, seed)
        d = oracle.descriptors(m.n, m.rowptr, m.col, m.val)
        assert d["psym"] == pytest.approx(scipy_psym(m), rel=0, abs=1e-15)
        deg = np.diff(m.rowptr).astype(np.float64)
        assert d["deg_cv"] == pytest.approx(deg.std() / deg.mean(), rel=1e-13)
    # hand-made: (0,1) stored twice summing to zero, (1,0) stored -> not
    # symmetric; (1,2)/(2,1) symmetric; an explicit zero at (2,2) is no entry
    m = ap.HostMatrix(3, np.array([0, 2, 4, 7], np.int32), np.array([1, 1, 0, 2, 1, 2, 0], np.int32),
                      np.array([0.5, -0.5, 2.0, 3.0, 4.0, 0.0, 1.0]))
    d = oracle.descriptors(m.n, m.rowptr, m.col, m.val)
    # P = {(1,0), (1,2), (2,1), (2,0)}; symmetric pairs: (1,2) and (2,1)
    assert d["psym"] == 2 / 4
    assert d["psym"] == scipy_psym(m)


def test_pick_ours_rule():
    # bench/dispatch.cu:14-17
    assert oracle.pick_ours(0.5, 53.0, 10) and not oracle.pick_ours(0.5, 54.0, 10)
    assert oracle.pick_ours(0.7, 47.0, 20165) and not oracle.pick_ours(0.7, 47.0, 20164)
    assert not oracle.pick_ours(0.7, 48.0, 10 ** 6)
