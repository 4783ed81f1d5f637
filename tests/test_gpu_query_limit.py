"""Queries too long for a workgroup's LDS (the query, its context tables and
the score matrix are held there; about 150K concatenated residues on MI355X)
are refused on the host with HB_ERR_UNSUPPORTED by every form, before
anything is queued or launched, and the requests submitted beside them are
unaffected: their HSPs equal the oracle's."""
import numpy as np
import pytest

import hbseed
import oracle_lib

pytestmark = pytest.mark.gpu

TOO_LONG = 200_000


@pytest.fixture(scope="module")
def setup():
    pq = hbseed.random_protein(5001, 600)
    res, offs = hbseed.make_db(5000, 3000, plant_query=pq, plant_frac=0.1)
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    short = []
    for i in range(4):
        q = hbseed.random_protein(5100 + i, 300 + 150 * i)
        q[:120] = pq[50 + 40 * i: 170 + 40 * i]
        short.append(hbseed.blastp_query(q))
    long_spec = hbseed.blastp_query(hbseed.random_protein(5200, TOO_LONG))
    exp = []
    for s in short:
        _, h = oracle_lib.OracleQuery.from_spec(s).scan(res, offs, threads=8)
        exp.append(h.view(np.int32).reshape(-1, 8))
    assert sum(len(e) for e in exp) > 0
    yield ctx, db, short, long_spec, exp
    db.close()
    ctx.close()


def _unsupported(fn):
    with pytest.raises(hbseed.HBError) as e:
        fn()
    assert e.value.code == -7, e.value


def test_plain_refuses_long_query(setup):
    ctx, db, short, long_spec, exp = setup
    _unsupported(lambda: hbseed.Query(ctx, long_spec))
    h = hbseed.search(ctx, db, short[0])[1].view(np.int32).reshape(-1, 8)
    assert np.array_equal(h, exp[0])


def test_batch_refuses_long_query_keeps_others(setup):
    ctx, db, short, long_spec, exp = setup
    b = hbseed.Batch(ctx, db)
    try:
        ids = [b.add(short[0]), b.add(short[1])]
        _unsupported(lambda: b.add(long_spec))
        ids += [b.add(short[2]), b.add(short[3])]
        b.run()
        for qid, e in zip(ids, exp):
            assert np.array_equal(b.results(qid).view(np.int32).reshape(-1, 8), e), qid
    finally:
        b.close()


def test_serve_refuses_long_query_keeps_others(setup):
    ctx, db, short, long_spec, exp = setup
    srv = hbseed.Server(ctx, db, linger_us=20000)
    try:
        tickets = [srv.submit(short[0]), srv.submit(short[1])]
        _unsupported(lambda: srv.submit(long_spec))
        tickets += [srv.submit(short[2]), srv.submit(short[3])]
        for t, e in zip(tickets, exp):
            assert srv.wait(t, 60000)
            assert np.array_equal(srv.results(t).view(np.int32).reshape(-1, 8), e), t
        st = srv.stats()
        assert st["submitted"] == 4 and st["completed"] == 4, st
    finally:
        srv.stop()
