"""Serving: jobs submitted while others run, small ones grouped into one
batch, admission against the HBM budget of every byte the slots hold (working
set bound before staging, exact C after the symbolic phase).  Each result must
be the same as the matrix computed alone (structure exact, values within
1e-12 * sum|a*b|), and the slots' device memory must stay within the budget."""
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
            # every job released: what stays charged is what the slots' pools cache
            assert st["bytes_reserved"] == st["device_bytes"]
            assert st["overruns"] == 0
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


def admitted_bytes(dev, op, slots):
    """Budget bytes one job is admitted for on a fresh server of `slots`
    slots (its working-set bound, bitmaps and C), and its result bytes."""
    with ap.Server(0, slots=slots, batch_max_jobs=1) as srv:
        j = srv.submit(dev, op)
        info = srv.wait(j)
        assert info["status_name"] == "APPT_OK", info
        st = srv.stats()
        assert st["overruns"] == 0 and st["peak_device_bytes"] <= st["peak_bytes_reserved"]
        srv.release(j)
        return st["peak_bytes_reserved"], info["reserved_bytes"]


def test_serve_waits_for_memory(eng):
    m = ap.stencil27(30, 0.0, 4)        # C ~ 27000 x 125 entries ~ 40 MB
    d = eng.upload(m)
    try:
        need, cbytes = admitted_bytes(d, "AA", 2)
        # room for one job, not for a second one beside a held result
        with ap.Server(0, slots=2, hbm_budget_bytes=need + cbytes // 2, batch_max_jobs=1,
                       admission_wait_ms=20000) as srv:
            say("case serve memory wait, budget", need + cbytes // 2)
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
            st = srv.stats()
            assert st["peak_device_bytes"] <= need + cbytes // 2 and st["overruns"] == 0
    finally:
        d.free()


def test_serve_tight_budget_four_slots(eng):
    # four slots under a budget of twice the largest single job's admission:
    # jobs (and groups) wait for one another's memory, none is refused, every
    # result equals the job alone, and the pools never hold more than the budget
    mats, ops = job_mix(7)
    dev = [eng.upload(m) for m in mats]
    try:
        alone = []
        for d, op in zip(dev, ops):
            r = eng.compute(d, op)
            try:
                alone.append(r.download())
            finally:
                r.release()
        budget = 2 * max(admitted_bytes(d, op, 1)[0] for d, op in zip(dev, ops))
        with ap.Server(0, slots=4, hbm_budget_bytes=budget, batch_row_limit=20000,
                       admission_wait_ms=60000) as srv:
            say("case serve tight budget: 40 jobs, 4 slots, budget", budget)
            ids = [srv.submit(d, op) for d, op in zip(dev, ops)]
            pending = dict(enumerate(ids))
            deadline = time.time() + 120
            while pending and time.time() < deadline:
                for k, j in list(pending.items()):
                    info = srv.poll(j)
                    if info["state"] != "done":
                        continue
                    assert info["status_name"] == "APPT_OK", (k, info)
                    got = srv.download(j)
                    np.testing.assert_array_equal(got.rowptr, alone[k].rowptr)
                    np.testing.assert_array_equal(got.col, alone[k].col)
                    assert np.all(np.abs(got.val - alone[k].val) <= TOL * ref(mats[k], ops[k])[3])
                    srv.release(j)
                    del pending[k]
                time.sleep(0.001)
            assert not pending, "jobs still outstanding"
            st = srv.stats()
            say("stats", st)
            assert st["completed"] == len(mats) and st["refused"] == 0 and st["failed"] == 0
            assert st["overruns"] == 0
            assert st["peak_device_bytes"] <= budget
            assert st["peak_bytes_reserved"] <= budget
    finally:
        for d in dev:
            d.free()
