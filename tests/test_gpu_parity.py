"""GPU parity: every product computed by the HIP engine (through the C ABI)
against the CPU oracle on the same inputs.  Structure must be identical;
values bit-equal where the row runs in a wave-level bin (C1, stencils,
banded, Erdos-Renyi), else within 1e-12 * sum|a*b| per entry."""
from pathlib import Path
import sys

import numpy as np
import pytest

import appt_spgemm_amd as ap
import oracle
from tests._util import assert_product, assert_structure, assert_values, ref, rows_sorted

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden" / "golden_small.npz"
OPS = ["AA", "AAt", "AAtS"]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = ap.Engine(0)
    yield e
    e.close()


def host(n, rp, col, val):
    return ap.HostMatrix(n, np.asarray(rp, np.int32), np.asarray(col, np.int32), np.asarray(val, np.float64))


@pytest.mark.parametrize("op", OPS)
def test_laplace_c1(eng, op):
    say("case laplace2d g=64", op)
    m = ap.laplace2d(64)
    got = eng.multiply(m, op)
    assert got.nnz == 51972
    assert_product(got, ref(m, op), exact=True)


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("drop", [0.0, 0.1])
def test_stencil_small(eng, op, drop):
    say("case stencil27 g=14", drop, op)
    m = ap.stencil27(14, drop, 5)
    got = eng.multiply(m, op)
    assert_product(got, ref(m, op), exact=True)


@pytest.mark.parametrize("op", OPS)
def test_golden_fixtures(eng, op):
    z = np.load(GOLD)
    for t in range(int(z["ncases"])):
        n = int(z[f"c{t}_n"])
        m = host(n, z[f"c{t}_rowptr"], z[f"c{t}_col"], z[f"c{t}_val"])
        say("case golden", t, op)
        got = eng.multiply(m, op)
        r = ref(m, op)
        # rows with duplicate columns may associate differently: tolerance only
        assert_product(got, r, exact=False)
        key = "AAt" if op == "AAtS" else op
        assert_structure(got.rowptr, got.col, z[f"c{t}_{key}_rowptr"], z[f"c{t}_{key}_col"])


@pytest.mark.parametrize("op", OPS)
def test_banded_and_er(eng, op):
    for m in (ap.banded(4000, 32, 0.7, 3), ap.banded(3000, 1, 0.7, 4), ap.erdos_renyi(6000, 12.0, 5),
              ap.erdos_renyi(5000, 2.0, 6)):
        say("case", m.meta, op)
        got = eng.multiply(m, op)
        assert_product(got, ref(m, op), exact=True)


def test_bin_report_narrow_span_rows(eng):
    """A half-bandwidth-32 band gives rows of about 2000-2800 products over a
    span of 5 words: past the last products class, yet each row stays in a
    64-lane wave window (no workgroup-per-row bin) and keeps its word list,
    so every numeric row runs in a list bin."""
    m = ap.banded(4000, 32, 0.7, 3)
    say("case bin report", m.meta)
    got = eng.multiply(m, "AA")
    assert_product(got, ref(m, "AA"), exact=True)
    br = eng.bin_report()
    say("bins", br)
    sym, num = br["symbolic"], br["numeric"]
    assert sum(v[0] for v in sym.values()) == m.n
    assert sum(v[0] for v in num.values()) == m.n
    assert "wg-window" not in sym and "wg-hash" not in sym
    assert sym["window/c2/G64"][1] > 2048
    assert all(k.startswith("list/") for k in num)


@pytest.mark.parametrize("op", OPS)
def test_rmat_small(eng, op):
    m = ap.rmat(12, 16, seed=1)
    say("case rmat scale 12", op)
    got = eng.multiply(m, op)
    assert_product(got, ref(m, op), exact=False)


@pytest.mark.parametrize("op", OPS)
def test_heavy_rows(eng, op):
    # arrow matrix: one dense row and column -> workgroup and heavy-row bins
    n = 20000
    rng = np.random.default_rng(7)
    rows = [np.arange(n)]
    cols = [np.arange(n)]
    rows.append(np.zeros(n - 1, np.int64)); cols.append(np.arange(1, n))
    rows.append(np.arange(1, n)); cols.append(np.zeros(n - 1, np.int64))
    for k in (300, 3000):
        r = rng.choice(n, k, replace=False); rows.append(r); cols.append(np.full(k, 17))
    r = np.concatenate(rows); c = np.concatenate(cols)
    key = np.unique(r * n + c)
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64); np.add.at(rp, r + 1, 1); rp = np.cumsum(rp)
    m = host(n, rp, c, rng.uniform(0, 1, len(c)))
    say("case arrow n=20000", op)
    got = eng.multiply(m, op)
    assert_product(got, ref(m, op), exact=False)


@pytest.fixture(scope="module")
def eng_heavy():
    # workgroup bins capped at 1 KB of LDS and 96-column windows: every row
    # wider than a few hundred columns takes the heavy-row bin, in many windows
    e = ap.Engine(0, tuning={"wg_lds_bytes": 1024, "heavy_window_cols": 96})
    yield e
    e.close()


def shuffled_rows(m, seed):
    """Same matrix with the entries of every row in random order."""
    rng = np.random.default_rng(seed)
    col = m.col.copy(); val = m.val.copy()
    for i in range(m.n):
        s, e = m.rowptr[i], m.rowptr[i + 1]
        p = s + rng.permutation(e - s)
        col[s:e] = m.col[p]; val[s:e] = m.val[p]
    return host(m.n, m.rowptr, col, val)


@pytest.mark.parametrize("op", OPS)
def test_heavy_windows(eng_heavy, op):
    # cursor mode (sorted B rows) over many windows
    for m in (ap.rmat(11, 16, seed=3), ap.erdos_renyi(3000, 40.0, 8), ap.banded(2500, 200, 0.5, 9)):
        say("case heavy windows", m.meta, op)
        got = eng_heavy.multiply(m, op)
        assert_product(got, ref(m, op), exact=False)


@pytest.mark.parametrize("op", OPS)
def test_heavy_windows_unsorted(eng_heavy, op):
    # rescan mode: B rows unsorted, every window filters by column
    m = shuffled_rows(ap.erdos_renyi(2000, 30.0, 10), 11)
    say("case heavy windows unsorted", op)
    got = eng_heavy.multiply(m, op)
    assert_product(got, ref(m, op), exact=False)


def random_rows(n, lens, col_hi, seed):
    """Rows of the given lengths with distinct random columns below col_hi."""
    rng = np.random.default_rng(seed)
    lens = np.asarray(lens, np.int64)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    parts = [np.sort(rng.choice(col_hi, size=int(lens[r]), replace=False)) for r in np.nonzero(lens)[0]]
    col = (np.concatenate(parts) if parts else np.zeros(0)).astype(np.int32)
    return host(n, rp.astype(np.int32), col, rng.uniform(-1, 1, size=len(col)))


def scattered_cases():
    # uniform random rows: outputs too scattered for a window (register-sort bins)
    yield "er_200k_16", ap.erdos_renyi(200_000, 16.0, 7)
    yield "er_60k_22", ap.erdos_renyi(60_000, 22.0, 8)   # rows on both sides of the 512-product limit
    # A entries whose B rows are empty: rows >= n/2 are empty, so about half
    # of every row's entries contribute no products (nA > products possible)
    n = 6000
    lens = np.where(np.arange(n) < n // 2, 12, 0)
    lens[7] = 300
    yield "empty_b_rows", random_rows(n, lens, n, 9)
    # A*A^T rows of ~700 and ~1400 products spread over 5M columns (sort
    # bins of 1024 and 2048 slots, 64-bit keys): 60000 rows scattered over
    # [0, n) with entries in the first 1000 columns
    n = 5_000_000
    rng = np.random.default_rng(11)
    lens = np.zeros(n, np.int64)
    rows = rng.choice(n, size=60_000, replace=False)
    lens[rows] = 2
    lens[rows[:300]] = 12
    lens[rows[300:400]] = 6
    yield "aat_wide_5m", random_rows(n, lens, 1000, 12)


@pytest.mark.parametrize("op", OPS)
def test_scattered_rows(eng, op):
    for name, m in scattered_cases():
        say("case scattered", name, op)
        got = eng.multiply(m, op)
        assert_product(got, ref(m, op), exact=True)


def test_scattered_rows_wide_columns(eng):
    # more than 2^23 columns: 64-bit sort keys
    m = ap.erdos_renyi(9_000_000, 1.5, 10)
    for op in ("AA", "AAt"):
        say("case scattered wide", op)
        got = eng.multiply(m, op)
        assert_product(got, ref(m, op), exact=True)


@pytest.fixture(scope="module")
def eng_radix():
    # every asymmetric transpose and every AAtS mirror through the radix
    # sort, however small
    e = ap.Engine(0, tuning={"transpose_radix_min_nnz": 1, "mirror_radix_min_nnz": 1})
    yield e
    e.close()


@pytest.mark.parametrize("op", ["AAt", "AAtS"])
def test_radix_transpose(eng_radix, op):
    cases = [(name, m) for name, m in edge_matrices()]
    cases += [("stencil_drop", ap.stencil27(14, 0.1, 5)), ("er_300", ap.erdos_renyi(300, 9.0, 4)),
              ("er_70000", ap.erdos_renyi(70_000, 3.0, 5)), ("banded", ap.banded(3000, 8, 0.7, 6)),
              ("rmat12", ap.rmat(12, 16, seed=3)), ("heavy_unsorted", shuffled_rows(ap.rmat(11, 16, seed=2), 3))]
    for name, m in cases:
        say("case radix transpose", name, op)
        got = eng_radix.multiply(m, op)
        exact = name not in ("duplicates", "rmat12", "heavy_unsorted")
        assert_product(got, ref(m, op), exact=exact)


def edge_matrices():
    yield "n0", host(0, [0], [], [])
    yield "n1", host(1, [0, 1], [0], [3.0])
    yield "n1_empty", host(1, [0, 0], [], [])
    yield "all_empty", host(10, [0] * 11, [], [])
    yield "empty_rows", host(6, [0, 2, 2, 3, 3, 5, 5], [1, 4, 0, 2, 5], [1.0, 2.0, 3.0, 4.0, 5.0])
    yield "unsorted", host(4, [0, 3, 4, 6, 7], [3, 0, 2, 1, 3, 0, 2], [1, 2, 3, 4, 5, 6, 7.0])
    yield "duplicates", host(3, [0, 3, 4, 6], [1, 1, 0, 2, 0, 0], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    yield "explicit_zeros", host(3, [0, 2, 3, 5], [0, 1, 2, 0, 2], [0.0, 1.0, 0.0, 2.0, 0.0])
    yield "cancel", host(2, [0, 2, 4], [0, 1, 0, 1], [1.0, 1.0, 1.0, -1.0])
    yield "last_col_only", host(5, [0, 1, 2, 3, 4, 5], [4, 4, 4, 4, 4], [1.0, 2.0, 3.0, 4.0, 5.0])


@pytest.mark.parametrize("op", OPS)
def test_edge_cases(eng, op):
    for name, m in edge_matrices():
        say("case edge", name, op)
        got = eng.multiply(m, op)
        assert_product(got, ref(m, op), exact=name != "duplicates")


def test_cancellation_structural(eng):
    m = dict(edge_matrices())["cancel"]
    got = eng.multiply(m, "AAt")
    assert got.rowptr.tolist() == [0, 2, 4] and got.col.tolist() == [0, 1, 0, 1]
    assert got.val.tolist() == [2.0, 0.0, 0.0, 2.0]


def drop_zeros(rp, col, val):
    """Numeric semantics from a structural product: scipy's eliminate_zeros."""
    import scipy.sparse as sp
    rp = np.asarray(rp, np.int64)
    M = sp.csr_matrix((np.asarray(val, np.float64), np.asarray(col), rp), shape=(len(rp) - 1, len(rp) - 1))
    M.eliminate_zeros()
    return M.indptr.astype(np.int64), M.indices, M.data


def small_int_values(m, seed, lo=-2, hi=3):
    """Same pattern, values drawn from small integers (products and sums are
    exact in FP64, so every cancellation is an exact zero in any order)."""
    v = np.random.default_rng(seed).integers(lo, hi, size=len(m.col)).astype(np.float64)
    return ap.HostMatrix(m.n, m.rowptr, m.col, v)


@pytest.mark.parametrize("op", OPS)
def test_eliminate_zeros(eng, op):
    cases = [("cancel", dict(edge_matrices())["cancel"]), ("explicit_zeros", dict(edge_matrices())["explicit_zeros"]),
             ("all_empty", dict(edge_matrices())["all_empty"]),
             ("stencil_int", small_int_values(ap.stencil27(12, 0.1, 3), 1, -1, 2)),
             ("er_int", small_int_values(ap.erdos_renyi(3000, 8.0, 2), 2)),
             ("rmat_int", small_int_values(ap.rmat(11, 16, seed=4), 3))]
    for name, m in cases:
        say("case eliminate_zeros", name, op)
        rp, col, val, _ = ref(m, op)
        erp, ecol, evals = drop_zeros(rp, col, val)
        got = eng.multiply(m, op, eliminate_zeros=True)
        assert_structure(got.rowptr, got.col, erp, ecol)
        np.testing.assert_array_equal(got.val, evals)
        assert not np.any(got.val == 0.0)
        if name == "stencil_int":
            assert len(ecol) < len(col)  # the case does cancel


def test_eliminate_zeros_cancel_2x2(eng):
    got = eng.multiply(dict(edge_matrices())["cancel"], "AAt", eliminate_zeros=True)
    assert got.rowptr.tolist() == [0, 1, 2] and got.col.tolist() == [0, 1] and got.val.tolist() == [2.0, 2.0]


def _poke_i32(addr, index, value):
    """Overwrite one int32 of device memory (hipMemcpy from the host)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    x = ctypes.c_int32(value)
    assert hip.hipMemcpy(ctypes.c_void_p(addr + 4 * index), ctypes.byref(x), ctypes.c_size_t(4), 1) == 0


def test_csr_check_device(eng):
    m = ap.stencil27(6, 0.1, 3)
    say("case csr_check")
    for poke, msg in ((None, None), (("col", 5, m.n), "column index"), (("col", 7, -1), "column index"),
                      (("rowptr", m.n, m.nnz + 1), "!= nnz"), (("rowptr", 3, 0), "decrease")):
        A = eng.upload(m)
        try:
            if poke is None:
                eng.check(A)
                continue
            which, idx, val = poke
            _poke_i32(A.view.col if which == "col" else A.view.rowptr, idx, val)
            with pytest.raises(ap.ApptError, match=msg):
                eng.check(A)
        finally:
            A.free()


def test_upload_rejects_invalid_host_csr(eng):
    bad = host(3, [0, 1, 2, 3], [0, 3, 1], [1.0, 2.0, 3.0])   # column 3 in a 3x3 matrix
    with pytest.raises(ap.ApptError, match="invalid"):
        eng.upload(bad)
    bad = host(3, [0, 2, 1, 3], [0, 1, 2], [1.0, 2.0, 3.0])   # decreasing row pointers
    with pytest.raises(ap.ApptError):
        eng.upload(bad)


def test_deterministic_repeat(eng):
    m = ap.stencil27(12, 0.1, 9)
    a = eng.multiply(m, "AAt")
    b = eng.multiply(m, "AAt")
    np.testing.assert_array_equal(a.col, b.col)
    np.testing.assert_array_equal(a.val, b.val)


def test_symmetric_invariants_gpu(eng):
    m = ap.laplace2d(40)
    aa, aat = eng.multiply(m, "AA"), eng.multiply(m, "AAt")
    np.testing.assert_array_equal(aa.rowptr, aat.rowptr)
    np.testing.assert_array_equal(aa.col, aat.col)
    np.testing.assert_array_equal(aa.val, aat.val)
    u = ap.stencil27(10, 0.1, 2)
    aats = eng.multiply(u, "AAtS")
    C = aats.to_scipy() if hasattr(aats, "to_scipy") else None
    C = ap.HostMatrix(u.n, aats.rowptr, aats.col, aats.val).to_scipy()
    assert (C != C.T).nnz == 0


def test_run_timed_checksums(eng):
    # run_ours contract: nnz + (s, as, sq) of the sorted values (src/spgemm_engine.cuh:194-196)
    m = ap.stencil27(16, 0.1, 4)
    A = eng.upload(m)
    try:
        for task in OPS:
            say("case run_timed", task)
            r = eng.run_timed(task, A, 3)
            rp, col, val, _ = ref(m, task)
            s, as_, sq = oracle.checksum(val)
            assert r.nnz == rp[-1]
            for a, b in ((r.s, s), (r.as_, as_), (r.sq, sq)):
                assert abs(a - b) <= 1e-6 * max(1.0, abs(b))  # bench/run_compare.py:42-43
            assert r.ms > 0
    finally:
        A.free()


def test_run_ours_free_function(eng):
    m = ap.laplace2d(32)
    A = eng.upload(m)
    try:
        r = ap.run_ours("AA", A.view.rowptr, A.view.col, A.view.val, m.n, m.nnz, 2)
        assert r.nnz == oracle.spgemm(m.n, m.rowptr, m.col, m.val, "AA")[0][-1]
    finally:
        A.free()


def test_descriptors_match_oracle(eng):
    # includes inputs with unsorted rows, duplicate positions, cancelling
    # duplicates and explicit zeros (psym over the merged numeric pattern)
    from tests._util import messy
    for m in (ap.rmat(11, 16, seed=3), ap.laplace2d(30), ap.stencil27(8, 0.1, 1), messy(2000, 1), messy(500, 2, 40)):
        A = eng.upload(m)
        try:
            for op in ("AA", "AAt"):
                say("case descriptors", m.n, m.nnz, op)
                d = eng.descriptors(A, op)
                o = oracle.descriptors(m.n, m.rowptr, m.col, m.val)
                for k in ("deg_mean", "deg_cv", "fill_aa", "fill_aat"):
                    assert d[k] == pytest.approx(o[k], rel=1e-13, abs=1e-15), k
                assert d["psym"] == o["psym"]      # same two integer counts
                assert d["products_aa"] == o["products_aa"]
                assert d["products_aat"] == o["products_aat"]
                assert d["bandwidth"] == o["bandwidth"]
                assert d["deg_max"] == o["deg_max"]
                fill = o["fill_aa"] if op == "AA" else o["fill_aat"]
                assert d["pick_ours"] == oracle.pick_ours(o["deg_cv"], fill, m.nnz)
        finally:
            A.free()


def test_phase_times_and_products(eng):
    m = ap.stencil27(20, 0.0, 1)
    eng.set_timing(True)
    try:
        A = eng.upload(m)
        R = eng.compute(A, "AA")
        t = eng.phase_times()
        assert t["products"] == (9 * 20 - 10) ** 3
        assert t["nnz_c"] == (5 * 20 - 6) ** 3
        assert t["total_ms"] > 0 and t["sync_points"] >= 2
        R.release()
        A.free()
    finally:
        eng.set_timing(False)


def test_budget_rejects_cleanly():
    e = ap.Engine(0, hbm_budget_bytes=1 << 20)
    try:
        m = ap.stencil27(30, 0.0, 1)
        A = e.upload(m)
        with pytest.raises(ap.ApptError) as ex:
            e.compute(A, "AA")
        assert ex.value.status_name == "APPT_ERR_OOM"
        A.free()
    finally:
        e.close()


def test_full_size_stencil_c2(eng):
    """27-point stencil 96^3 (n = 884,736), A*A, A*A^T and AAtS: closed-form
    counts, sorted rows, row sums, and a sample of rows bit-equal to the
    oracle."""
    for drop, seed in ((0.0, 1), (0.1, 1)):
        m = ap.stencil27(96, drop, seed)
        say("case stencil27 g=96 drop", drop)
        A = eng.upload(m)
        try:
            for op in ("AA", "AAt", "AAtS"):
                R = eng.compute(A, op)
                got = R.download()
                R.release()
                if drop == 0.0:
                    assert got.nnz == 474 ** 3
                assert rows_sorted(got.rowptr, got.col)
                S = m.to_scipy()
                B = S if op == "AA" else S.T.tocsr()
                C = got.to_scipy() if False else ap.HostMatrix(m.n, got.rowptr, got.col, got.val).to_scipy()
                ones = np.ones(m.n)
                np.testing.assert_allclose(C @ ones, S @ (B @ ones), rtol=1e-9, atol=1e-9)
                # AAtS: the mirrored upper triangle; with strictly sorted rows
                # c_ij and c_ji sum the same products in the same order, so the
                # A*A^T rows of the oracle are the reference bit for bit
                oop = "AAt" if op == "AAtS" else op
                for r0 in (0, 400000, m.n - 3000):
                    off, c2, v2 = oracle.rows_range(m.n, m.rowptr, m.col, m.val, r0, r0 + 3000, oop)
                    np.testing.assert_array_equal(got.rowptr[r0:r0 + 3001] - got.rowptr[r0], off)
                    lo, hi = got.rowptr[r0], got.rowptr[r0 + 3000]
                    np.testing.assert_array_equal(got.col[lo:hi], c2)
                    np.testing.assert_array_equal(got.val[lo:hi], v2)
        finally:
            A.free()


@pytest.mark.parametrize("op", ["AA", "AAt"])
def test_row_panels(eng, op):
    # symbolic row pointers, then C in row panels of bounded bytes
    m = ap.rmat(12, 16, seed=2)
    say("case row panels rmat12", op)
    r = ref(m, op)
    A = eng.upload(m)
    try:
        rp = eng.symbolic_rowptr(A, op)
        np.testing.assert_array_equal(rp, r[0])
        parts = ap.panels(rp, max_bytes=400_000)
        assert len(parts) > 3
        cols, vals, rps = [], [], [np.zeros(1, np.int64)]
        for r0, r1 in parts:
            R = eng.compute_rows(A, op, r0, r1)
            try:
                h = R.download()
            finally:
                R.release()
            assert h.n == r1 - r0
            cols.append(h.col); vals.append(h.val)
            rps.append(np.asarray(h.rowptr[1:], np.int64) + rps[-1][-1])
        got = ap.HostMatrix(m.n, np.concatenate(rps), np.concatenate(cols), np.concatenate(vals))
        assert_product(got, r, exact=False)
    finally:
        A.free()


@pytest.mark.parametrize("op", ["AA", "AAt"])
def test_row_panels_planned_heavy_rows(eng_heavy, op):
    # heavy rows of panels take their counts from the row-pointer plan of the
    # same view; a plan-free engine computes the same panels from scratch.
    # Identity plus a dense row 0 and column-0 entries in 40 rows: those rows'
    # outputs span all n columns, past this engine's workgroup bins (heavy)
    n = 12000
    rng = np.random.default_rng(12)
    some = rng.choice(np.arange(1, n), 40, replace=False)
    r = np.concatenate([np.arange(n), np.zeros(n - 1, np.int64), some])
    c = np.concatenate([np.arange(n), np.arange(1, n), np.zeros(40, np.int64)])
    key = np.unique(r * n + c)
    r, c = key // n, key % n
    rp = np.zeros(n + 1, np.int64); np.add.at(rp, r + 1, 1); rp = np.cumsum(rp)
    m = host(n, rp, c, rng.uniform(0, 1, len(c)))
    say("case planned heavy panels n=12000", op)
    want = ref(m, op)
    A = eng_heavy.upload(m)
    fresh = ap.Engine(0, tuning={"wg_lds_bytes": 1024, "heavy_window_cols": 96})
    B = fresh.upload(m)
    try:
        rp = eng_heavy.symbolic_rowptr(A, op)
        np.testing.assert_array_equal(rp, want[0])
        # about five panels, so plan rows are taken at non-zero panel offsets
        parts = ap.panels(rp, max_bytes=(12 * int(rp[-1]) + 8 * (n + 1)) // 5)
        assert len(parts) > 1
        cols, vals, rps = [], [], [np.zeros(1, np.int64)]
        heavy_rows = 0
        for r0, r1 in parts:
            R = eng_heavy.compute_rows(A, op, r0, r1)
            heavy_rows += eng_heavy.bin_report()["symbolic"].get("heavy", (0, 0))[0]
            F = fresh.compute_rows(B, op, r0, r1)
            try:
                h, f = R.download(), F.download()
            finally:
                R.release()
                F.release()
            np.testing.assert_array_equal(h.rowptr, f.rowptr)
            np.testing.assert_array_equal(h.col, f.col)
            cols.append(h.col); vals.append(h.val)
            rps.append(np.asarray(h.rowptr[1:], np.int64) + rps[-1][-1])
        assert heavy_rows > 0
        got = ap.HostMatrix(m.n, np.concatenate(rps), np.concatenate(cols), np.concatenate(vals))
        assert_product(got, want, exact=False)
    finally:
        A.free()
        B.free()
        fresh.close()
