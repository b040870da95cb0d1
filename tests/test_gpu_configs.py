"""The measured configurations at their full sizes, checked against the
oracle (the methodology of bench/run_compare.py:80-86: structure equal,
values within tolerance):

* R-MAT 2^21 (edge factor 16, seed 1), A*A and A*A^T: the symbolic row
  pointers, then row panels (the first one holds the hub rows, the last one
  the tail); every panel's nnz equals the plan; sampled rows, including the
  64 heaviest, equal the oracle's.
* The 4096-matrix uneven batch, every matrix against the oracle.
* A Poisson stream of served jobs from the Laplacian / stencil / R-MAT
  generators with n up to 2e6 under the 200 GB budget: every result equals
  the same job run alone and the oracle on sampled rows; jobs whose result
  cannot fit the budget are refused."""
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import appt_spgemm_amd as ap
import oracle
from appt_spgemm_amd.workloads import batch_specs, generate, serve_job, truncate
from tests._util import TOL, assert_structure, assert_values, ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = ap.Engine(0)
    yield e
    e.close()


def row_products(m, op):
    rp = np.asarray(m.rowptr, np.int64)
    col = np.asarray(m.col, np.int64)
    blen = np.diff(rp) if op == "AA" else np.bincount(col, minlength=m.n)
    per_entry = blen[col]
    return np.add.reduceat(per_entry, rp[:-1].clip(max=max(len(col) - 1, 0))) * (np.diff(rp) > 0)


def oracle_rows(m, b, rows, op, workers=16):
    """Oracle (rowptr, col, val, scale) of single rows, in parallel."""
    def one(i):
        return oracle.rows_range(m.n, m.rowptr, m.col, m.val, int(i), int(i) + 1, op, threads=1, b=b,
                                 want_scale=True)
    with ThreadPoolExecutor(workers) as ex:
        return dict(zip(rows, ex.map(one, rows)))


@pytest.fixture(scope="module")
def rmat21():
    t = time.time()
    m = ap.rmat(21, 16, seed=1)
    say(f"rmat 2^21: n {m.n} nnz {m.nnz} generated in {time.time() - t:.1f} s")
    return m


@pytest.mark.parametrize("op", ["AA", "AAt"])
def test_c3_rmat21_panels(eng, rmat21, op):
    m = rmat21
    rng = np.random.default_rng(5)
    f = row_products(m, op)
    heavy = np.argsort(f)[-64:]
    sample = np.unique(np.concatenate([heavy, rng.choice(m.n, 128, replace=False), [0, m.n - 1]]))
    b = (m.rowptr, m.col, m.val) if op == "AA" else oracle.transpose(m.n, m.rowptr, m.col, m.val)
    t = time.time()
    want = oracle_rows(m, b, sample, op)
    say(f"case c3 {op}: oracle on {len(sample)} rows in {time.time() - t:.1f} s")
    A = eng.upload(m)
    try:
        rp = eng.symbolic_rowptr(A, op)
        say(f"  plan: nnz(C) = {rp[-1]}")
        for i in sample:
            assert rp[i + 1] - rp[i] == want[i][0][-1], f"row {i}: count {rp[i + 1] - rp[i]} vs {want[i][0][-1]}"
        parts = ap.panels(rp, max_bytes=40 * 10 ** 9)
        assert len(parts) > 2
        first, last = parts[0], parts[-1]
        chosen = [first, last]
        for r0, r1 in parts[1:-1]:        # any other panel holding one of the heaviest rows
            if np.any((heavy >= r0) & (heavy < r1)):
                chosen.append((r0, r1))
        say(f"  {len(parts)} panels; checking {chosen}")
        checked = 0
        for r0, r1 in chosen:
            t = time.time()
            R = eng.compute_rows(A, op, r0, r1)
            try:
                assert R.nnz == rp[r1] - rp[r0]
                assert R.res.rowptr_is_64 == (1 if R.nnz > 2 ** 31 - 1 else 0)
                inside = sample[(sample >= r0) & (sample < r1)]
                for i in inside:
                    h = R.download_rows(int(i) - r0, int(i) - r0 + 1)
                    o_rp, o_col, o_val, o_scale = want[i]
                    assert_structure(h.rowptr, h.col, o_rp, o_col)
                    assert_values(h.val, o_val, o_scale, exact=False)
                    checked += 1
            finally:
                R.release()
            say(f"  panel [{r0}, {r1}) nnz {rp[r1] - rp[r0]}: {len(inside)} sampled rows equal ({time.time() - t:.1f} s)")
        assert checked >= 64
    finally:
        A.free()


def test_c4_full_batch(eng):
    specs, ops = batch_specs(4096, 1)
    t = time.time()
    mats = generate(specs)
    say(f"case c4: 4096 matrices, sum n {sum(m.n for m in mats)}, sum nnz {sum(m.nnz for m in mats)}, "
        f"generated in {time.time() - t:.1f} s")
    dev = [eng.upload(m) for m in mats]
    try:
        with ap.Batch(eng, [m.n for m in mats], [m.nnz for m in mats], ops) as b:
            b.gather(dev)
            b.plan(True)
            b.execute()
            info = b.info()
            say(f"  batch: nnz(C) {info['nnz_c']}, products {info['products']}, graph {info['graph']}")

            def check(k):
                got = b.download(k)
                rp, col, val, scale = oracle.spgemm(mats[k].n, mats[k].rowptr, mats[k].col, mats[k].val, ops[k],
                                                    threads=1)
                assert_structure(got.rowptr, got.col, rp, col)
                assert_values(got.val, val, scale, exact=False)
                return int(np.count_nonzero(got.val == val)), len(val)

            t = time.time()
            with ThreadPoolExecutor(16) as ex:
                res = list(ex.map(check, range(len(mats))))
            same = sum(r[0] for r in res)
            total = sum(r[1] for r in res)
            say(f"  4096 matrices equal the oracle ({time.time() - t:.1f} s); bit-equal values {same} of {total}")
            assert total == info["nnz_c"]
    finally:
        for d in dev:
            d.free()


def test_c5_served_poisson_stream(eng):
    """>= 200 jobs, Poisson arrivals, 4 slots, 200 GB budget."""
    pool_n, jobs_n, qps = 24, 240, 150.0
    pool = generate([lambda r=np.random.default_rng(7919 + k), k=k: serve_job(r, 1000, 2_000_000, seed=k)
                     for k in range(pool_n)], workers=8)
    # and the largest R-MAT job of the range (2^21 rows cut to n = 2e6): its
    # result, about 300 GB, can never fit the 200 GB budget
    pool.append((truncate(ap.rmat(21, 16, seed=11), 2_000_000), "AA", "rmat"))
    pool_n += 1
    say("case c5: pool n =", [m.n for m, _, _ in pool])
    dev = [eng.upload(m) for m, _, _ in pool]
    budget = 200 * 10 ** 9
    rng = np.random.default_rng(3)
    picks = rng.integers(0, pool_n, jobs_n)
    picks[rng.choice(jobs_n, 4, replace=False)] = pool_n - 1
    ops = [pool[k][1] if rng.random() < 0.5 else ("AAt" if pool[k][1] == "AA" else "AA") for k in picks]
    # what each (matrix, op) gives alone: its size, digests, and sampled rows
    alone = {}
    try:
        for k, op in sorted(set(zip(picks.tolist(), ops))):
            m, _, fam = pool[k]
            rp = eng.symbolic_rowptr(dev[k], op)
            cbytes = (8 if rp[-1] > 2 ** 31 - 1 else 4) * (m.n + 1) + 12 * int(rp[-1])
            entry = {"bytes": cbytes, "fam": fam, "nnz": int(rp[-1])}
            if cbytes + (64 * m.n + 28 * m.nnz) < budget:
                R = eng.compute(dev[k], op)
                try:
                    entry["digest"] = R.digest()
                    r0 = int(rng.integers(0, max(1, m.n - 16)))
                    entry["rows"] = (r0, R.download_rows(r0, min(m.n, r0 + 16)))
                finally:
                    R.release()
                eng.trim()
            alone[(k, op)] = entry
        say("  alone:", {f"{k}/{op}": (v["fam"], v["nnz"]) for (k, op), v in alone.items()})
        with ap.Server(0, slots=4, hbm_budget_bytes=budget) as srv:
            ids = []
            t0 = time.perf_counter()
            t_next = t0
            for k, op in zip(picks, ops):
                t_next += rng.exponential(1.0 / qps)
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                ids.append(srv.submit(dev[k], op))
            say(f"  submitted {jobs_n} jobs in {time.perf_counter() - t0:.2f} s")
            refused = ok = 0
            for j, k, op in zip(ids, picks, ops):
                info = srv.wait(j)
                a = alone[(int(k), op)]
                if "digest" not in a:
                    assert info["status_name"] == "APPT_ERR_ADMISSION", info
                    refused += 1
                    continue
                assert info["status_name"] == "APPT_OK", (k, op, info)
                C = srv.result(j)
                ds, dv = C.digest(eng)
                assert C.nnz == a["nnz"] and ds == a["digest"][0], (k, op)
                if a["fam"] != "rmat":       # wave-level bins only: values bit-identical
                    assert dv == a["digest"][1], (k, op)
                r0, ha = a["rows"]
                hs = C.download_rows(r0, r0 + ha.n)
                np.testing.assert_array_equal(hs.col, ha.col)
                m = pool[int(k)][0]
                o_rp, o_col, o_val, o_scale = oracle.rows_range(m.n, m.rowptr, m.col, m.val, r0, r0 + ha.n, op,
                                                                want_scale=True)
                assert_structure(hs.rowptr, hs.col, o_rp, o_col)
                assert_values(hs.val, o_val, o_scale, exact=False)
                assert np.all(np.abs(hs.val - ha.val) <= TOL * o_scale)
                srv.release(j)
                ok += 1
            st = srv.stats()
        say(f"  served {ok}, refused {refused}; p50 {st['p50_ms']:.2f} ms p99 {st['p99_ms']:.2f} ms, "
            f"{st['jobs_per_s']:.1f} jobs/s, peak reserved {st['peak_bytes_reserved'] / 1e9:.1f} GB")
        assert ok + refused == jobs_n and st["completed"] == ok and st["refused"] == refused
        assert refused >= 4
        assert st["peak_bytes_reserved"] <= budget and st["peak_device_bytes"] <= budget
        assert st["overruns"] == 0
        # every job released: what stays charged is what the slots' pools cache
        assert st["bytes_reserved"] == st["device_bytes"]
    finally:
        for d in dev:
            d.free()


@pytest.mark.parametrize("op", ["AA", "AAt"])
def test_rmat18_heavy_windows_default_shapes(eng, op):
    """R-MAT 2^18 in one call with the default heavy-row windows: rows wider
    than eight dense windows take sparse windows shaped per row (held values,
    cut windows, resident entries), the widest hubs (several 1024-entry
    chunks) dense windows; the 64 heaviest rows and 128 random ones equal the
    oracle's."""
    m = ap.rmat(18, 16, seed=4)
    rng = np.random.default_rng(9)
    f = row_products(m, op)
    heavy = np.argsort(f)[-64:]
    sample = np.unique(np.concatenate([heavy, rng.choice(m.n, 128, replace=False), [0, m.n - 1]]))
    b = (m.rowptr, m.col, m.val) if op == "AA" else oracle.transpose(m.n, m.rowptr, m.col, m.val)
    want = oracle_rows(m, b, sample, op)
    say(f"case rmat 2^18 {op}: max row {int(np.diff(m.rowptr).max())} entries, {len(sample)} sampled rows")
    A = eng.upload(m)
    try:
        R = eng.compute(A, op)
        try:
            rows = eng.bin_report()["numeric"].get("heavy", (0, 0))[0]
            say(f"  nnz(C) {R.nnz}, heavy rows {rows}")
            assert rows > 1000
            for i in sample:
                h = R.download_rows(int(i), int(i) + 1)
                o_rp, o_col, o_val, o_scale = want[i]
                assert_structure(h.rowptr, h.col, o_rp, o_col)
                assert_values(h.val, o_val, o_scale, exact=False)
        finally:
            R.release()
    finally:
        A.free()
