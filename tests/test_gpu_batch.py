"""Batched path: many matrices as one block-diagonal product, planned once and
replayed from a HIP graph.  Every result must equal the oracle (bit-exact on
these banded / Erdos-Renyi inputs, which run in wave-level bins) and the
same matrix computed alone."""
import sys

import numpy as np
import pytest

import appt_spgemm_amd as ap
from tests._util import TOL, assert_product, ref

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def mixed(count, seed, nmax=3000):
    """Uneven mix: n log-uniform, alternating banded / Erdos-Renyi, random op."""
    rng = np.random.default_rng(seed)
    mats, ops = [], []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(16), np.log(nmax))))
        if k % 2 == 0:
            m = ap.banded(n, int(rng.integers(1, 33)), 0.7, seed * 1000 + k)
        else:
            m = ap.erdos_renyi(n, float(rng.uniform(2, 20)), seed * 1000 + k)
        mats.append(m)
        ops.append("AA" if rng.random() < 0.5 else "AAt")
    return mats, ops


@pytest.fixture(scope="module")
def eng():
    e = ap.Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("graph", [True, False])
def test_batch_matches_oracle(eng, graph):
    mats, ops = mixed(24, 1)
    say("case batch 24 graph", graph)
    with ap.Batch.from_host(eng, mats, ops) as b:
        b.plan(graph)
        info = b.info()
        assert info["graph"] == (1 if graph else 0)
        for rep in range(3):
            b.execute()
            for k, (m, op) in enumerate(zip(mats, ops)):
                assert_product(b.download(k), ref(m, op), exact=True)
        assert info["nnz_c"] == sum(ref(m, op)[0][-1] for m, op in zip(mats, ops))


def test_batch_alone_equals_mixed(eng):
    mats, ops = mixed(16, 2)
    say("case batch alone vs mixed")
    with ap.Batch.from_host(eng, mats, ops) as b:
        b.plan(True)
        b.execute()
        for k, (m, op) in enumerate(zip(mats, ops)):
            alone = eng.multiply(m, op)
            got = b.download(k)
            np.testing.assert_array_equal(got.rowptr, alone.rowptr)
            np.testing.assert_array_equal(got.col, alone.col)
            # same bins, same accumulation order where a row is owned by one wave
            scale = ref(m, op)[3]
            assert np.all(np.abs(got.val - alone.val) <= TOL * scale)


def test_batch_new_values_replay(eng):
    # the structure is fixed at plan time; the graph picks up new values
    mats, ops = mixed(10, 3)
    say("case batch new values")
    with ap.Batch.from_host(eng, mats, ops) as b:
        b.plan(True)
        b.execute()
        rng = np.random.default_rng(9)
        mats2 = [ap.HostMatrix(m.n, m.rowptr, m.col, rng.uniform(-1, 1, m.nnz)) for m in mats]
        b.upload(ap.pack(mats2))
        b.execute()
        for k, (m, op) in enumerate(zip(mats2, ops)):
            assert_product(b.download(k), ref(m, op), exact=True)


def test_batch_gather_and_edges(eng):
    # device-side gather; includes an empty matrix, a matrix with no entries
    # and a 1x1 matrix
    mats, ops = mixed(6, 4, nmax=500)
    mats += [ap.HostMatrix(0, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0)),
             ap.HostMatrix(5, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0)),
             ap.HostMatrix(1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([3.0]))]
    ops += ["AAt", "AA", "AAt"]
    say("case batch gather + edges")
    dev = [eng.upload(m) for m in mats]
    try:
        with ap.Batch(eng, [m.n for m in mats], [m.nnz for m in mats], ops) as b:
            b.gather(dev)
            b.plan(True)
            b.execute()
            for k, (m, op) in enumerate(zip(mats, ops)):
                assert_product(b.download(k), ref(m, op), exact=True)
    finally:
        for d in dev:
            d.free()


def test_batch_transpose_paths(eng):
    # A*A^T blocks of at most 16384 rows are transposed one workgroup each
    # (LDS column counts); a larger one sends the whole batch to the global
    # counts-and-cursors transpose.  Both must give the oracle's products.
    small, sops = mixed(8, 6, nmax=2000)
    for extra in (ap.banded(16384, 9, 0.7, 61), ap.erdos_renyi(20000, 3.0, 62)):
        mats, ops = small + [extra], sops + ["AAt"]
        say("case batch transpose path, largest A*A^T n =", extra.n)
        with ap.Batch.from_host(eng, mats, ops) as b:
            b.plan(True)
            b.execute()
            for k, (m, op) in enumerate(zip(mats, ops)):
                assert_product(b.download(k), ref(m, op), exact=True)


def test_batch_autotune(eng):
    mats, ops = mixed(12, 5)
    say("case batch autotune")
    with ap.Batch.from_host(eng, mats, ops) as b:
        b.plan(True)
        t = b.autotune(1)
        assert len(t["f_class_max"]) == 3 and len(t["group_for_len"]) == 3
        assert b.info()["tuning"] == t
        b.execute()
        for k, (m, op) in enumerate(zip(mats, ops)):
            assert_product(b.download(k), ref(m, op), exact=True)


def test_batch_rejects_aats(eng):
    with pytest.raises(ap.ApptError):
        ap.Batch(eng, [4], [4], ["AAtS"])
