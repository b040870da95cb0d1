"""Serving: jobs submitted while others run, small ones grouped into one
batch, admission by exact result bytes against the HBM budget.  Each result
must be the same as the matrix computed alone (structure exact, values within
1e-12 * sum|a*b|)."""
import sys
import threading
import time

import numpy as np
import pytest

import appt_spgemm_amd as ap
from tests._util import TOL, assert_product, ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = ap.Engine(0)
    yield e
    e.close()


def job_mix(seed):
    rng = np.random.default_rng(seed)
    mats, ops = [], []
    for k in range(40):
        kind = k % 5
        if kind == 0:
            m = ap.stencil27(int(rng.integers(8, 40)), 0.1 if rng.random() < 0.5 else 0.0, k)
        elif kind == 1:
            m = ap.laplace2d(int(rng.integers(16, 300)))
        elif kind == 2:
            m = ap.rmat(int(rng.integers(8, 13)), 16, seed=k)
        elif kind == 3:
            m = ap.banded(int(rng.integers(50, 5000)), int(rng.integers(1, 33)), 0.7, k)
        else:
            m = ap.erdos_renyi(int(rng.integers(50, 5000)), float(rng.uniform(2, 20)), k)
        mats.append(m)
        ops.append("AA" if rng.random() < 0.5 else "AAt")
    return mats, ops


def test_serve_mix_equals_alone(eng):
    mats, ops = job_mix(1)
    dev = [eng.upload(m) for m in mats]
    try:
        with ap.Server(0, slots=3, batch_row_limit=20000) as srv:
            say("case serve mix: submit 40 jobs, staggered")
            ids = []
            for k, (d, op) in enumerate(zip(dev, ops)):
                ids.append(srv.submit(d, op))
                if k % 7 == 3:
                    time.sleep(0.002)  # arrivals while others are on the GPU
            for j, m, op, d in zip(ids, mats, ops, dev):
                info = srv.wait(j)
                assert info["status_name"] == "APPT_OK", info
                got = srv.download(j)
                r = ref(m, op)
                assert_product(got, r, exact=False)
                alone = eng.compute(d, op)
                try:
                    a = alone.download()
                finally:
                    alone.release()
                np.testing.assert_array_equal(got.rowptr, a.rowptr)
                np.testing.assert_array_equal(got.col, a.col)
                assert np.all(np.abs(got.val - a.val) <= TOL * r[3])
                srv.release(j)
            st = srv.stats()
            say("stats", st)
            assert st["completed"] == len(mats) and st["refused"] == 0 and st["failed"] == 0
            assert st["groups_run"] >= 1
            assert st["p99_ms"] >= st["p50_ms"] > 0
            assert st["bytes_reserved"] == 0
    finally:
        for d in dev:
            d.free()


def test_serve_admission_refuses_oversized(eng):
    big = ap.stencil27(40, 0.0, 3)      # C: 64000 rows x ~125 -> ~96 MB
    small = ap.laplace2d(32)
    db, ds = eng.upload(big), eng.upload(small)
    try:
        with ap.Server(0, slots=2, hbm_budget_bytes=48 << 20, admission_wait_ms=200) as srv:
            say("case serve admission")
            jb = srv.submit(db, "AA")
            js = srv.submit(ds, "AA")
            ib, is_ = srv.wait(jb), srv.wait(js)
            assert ib["status_name"] == "APPT_ERR_ADMISSION"
            assert is_["status_name"] == "APPT_OK"
            assert_product(srv.download(js), ref(small, "AA"), exact=True)
            srv.release(js)
            st = srv.stats()
            assert st["refused"] == 1 and st["completed"] == 1
    finally:
        db.free()
        ds.free()


def test_serve_waits_for_memory(eng):
    m = ap.stencil27(30, 0.0, 4)        # C ~ 27000 x 125 entries ~ 40 MB
    d = eng.upload(m)
    try:
        need = 4 * (m.n + 1) + 12 * int(ref(m, "AA")[0][-1])
        work = 64 * m.n + 28 * m.nnz + (8 << 20)   # the server's working-set estimate
        # room for one job (result + working set) beside nothing, not beside a held result
        with ap.Server(0, slots=2, hbm_budget_bytes=need + work + need // 2, batch_max_jobs=1,
                       admission_wait_ms=20000) as srv:
            say("case serve memory wait")
            j1 = srv.submit(d, "AA")
            assert srv.wait(j1)["status_name"] == "APPT_OK"
            j2 = srv.submit(d, "AA")
            time.sleep(0.3)
            assert srv.poll(j2)["state"] in ("queued", "running")   # waiting for j1's memory
            t = threading.Timer(0.05, srv.release, args=(j1,))
            t.start()
            info = srv.wait(j2)
            t.join()
            assert info["status_name"] == "APPT_OK" and info["latency_ms"] >= 300
            assert_product(srv.download(j2), ref(m, "AA"), exact=True)
            srv.release(j2)
    finally:
        d.free()
