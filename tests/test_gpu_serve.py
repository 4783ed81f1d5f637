"""Served form: requests submitted from several threads at staggered times,
mixed BLASTP/BLASTX and lengths, folded into concurrent batched scans; each
request's HSPs equal the query searched alone through the plain path and the
oracle (bit-exact), also when the memory budget limits the batches and when
a batch's output overflows its reservation and is rerun."""
import threading
import time

import numpy as np
import pytest

import hbseed
import oracle_lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    pq = hbseed.random_protein(4001, 700)
    res, offs = hbseed.make_db(4000, 8000, plant_query=pq, plant_frac=0.1)
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    specs = []
    lens = hbseed.lognormal_lengths(4002, 20, 350.0, 0.7, 50, 3000)
    for i, L in enumerate(lens):
        q = hbseed.random_protein(4100 + i, int(L))
        if i % 4 == 0:
            n = min(150, len(q))
            q[:n] = pq[300:300 + n]
        specs.append(hbseed.blastp_query(q))
    for i in range(5):   # 20% BLASTX, 600-3000 nt
        seg = hbseed.homolog(4300 + i, pq[100 + 60 * i: 200 + 60 * i], 0.7, 1)
        specs.append(hbseed.blastx_query(hbseed.blastx_nt(4310 + i, 600 + 500 * i, seg,
                                                          seg[::-1].copy())))
    order = np.random.default_rng(4400).permutation(len(specs))
    specs = [specs[i] for i in order]
    alone = [hbseed.search(ctx, db, s)[1].view(np.int32).reshape(-1, 8).copy() for s in specs]
    yield ctx, db, res, offs, specs, alone
    db.close()
    ctx.close()


def _submit_all(srv, specs, n_threads=3, seed=4500):
    """Submits specs from n_threads threads with exponential gaps; returns
    tickets in spec order."""
    tickets = [None] * len(specs)
    errors = []

    def worker(t):
        rng = np.random.default_rng(seed + t)
        for i in range(t, len(specs), n_threads):
            time.sleep(float(rng.exponential(0.004)))
            try:
                tickets[i] = srv.submit(specs[i])
            except Exception as e:   # reported below
                errors.append(e)

    th = [threading.Thread(target=worker, args=(t,)) for t in range(n_threads)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    assert not errors, errors
    return tickets


def _check(srv, tickets, alone):
    for i, t in enumerate(tickets):
        assert srv.wait(t, 60000)
        got = srv.results(t).view(np.int32).reshape(-1, 8)
        assert np.array_equal(got, alone[i]), (i, len(got), len(alone[i]))


def test_alone_equals_plain_and_oracle(setup):
    ctx, db, res, offs, specs, alone = setup
    assert sum(len(a) for a in alone) > 0
    for i in (0, 5, len(specs) - 1):
        _, h = oracle_lib.OracleQuery.from_spec(specs[i]).scan(res, offs, threads=8)
        assert np.array_equal(alone[i], h.view(np.int32).reshape(-1, 8))


@pytest.mark.parametrize("max_inflight,max_batch_queries", [(2, 4), (3, 64), (1, 1)])
def test_serve_mix_equals_alone(setup, max_inflight, max_batch_queries):
    ctx, db, res, offs, specs, alone = setup
    with hbseed.Server(ctx, db, max_inflight=max_inflight,
                       max_batch_queries=max_batch_queries) as srv:
        tickets = _submit_all(srv, specs)
        _check(srv, tickets, alone)
        st = srv.stats()
    assert st["completed"] == len(specs) and st["rejected"] == 0
    assert st["batches"] >= -(-len(specs) // max_batch_queries)
    assert st["peak_inflight"] <= max_inflight
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]


def test_serve_memory_budget(setup):
    ctx, db, res, offs, specs, alone = setup
    budget = 3 << 20   # a few typical requests
    with hbseed.Server(ctx, db, mem_budget=budget, max_inflight=2) as srv:
        tickets = _submit_all(srv, specs, n_threads=2)
        _check(srv, tickets, alone)
        st = srv.stats()
        assert st["peak_bytes"] <= budget
        assert st["batches"] > len(specs) // 4
        with pytest.raises(hbseed.HBError) as e:   # one request above the budget
            srv.submit(hbseed.blastp_query(hbseed.random_protein(4600, 40000)))
        assert e.value.code == -4


def test_serve_output_overflow_rerun(setup):
    ctx, db, res, offs, specs, alone = setup
    # one batch of every request (held up to 2 s for the batch to fill) with
    # no output reservation beyond 256 records per request: its records
    # overflow the buffer and the pass is rerun with the counted size
    assert sum(len(a) for a in alone) > 256 * len(specs) + 4096
    with hbseed.Server(ctx, db, hsp_per_cell=0.0, max_batch_queries=len(specs),
                       max_batch_residues=1 << 30, linger_us=2_000_000, max_inflight=1) as srv:
        tickets = _submit_all(srv, specs, n_threads=1)
        _check(srv, tickets, alone)
        st = srv.stats()
    assert st["batches"] == 1 and st["reruns"] == 1


def test_serve_rejects_bad_residue(setup):
    ctx, db, res, offs, specs, alone = setup
    bad = hbseed.blastp_query(hbseed.random_protein(4700, 100))
    bad.seq[10] = 40
    with hbseed.Server(ctx, db) as srv:
        with pytest.raises(hbseed.HBError) as e:
            srv.submit(bad)
        assert e.value.code == -5
        t = srv.submit(specs[0])
        assert srv.wait(t, 60000)
        assert np.array_equal(srv.results(t).view(np.int32).reshape(-1, 8), alone[0])
