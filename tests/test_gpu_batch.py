"""Batched form: every query's HSPs from one batched scan equal the oracle's
for that query alone (bit-exact), across groups, size classes, BLASTP and
BLASTX, graph replay and the autotuned group size."""
import numpy as np
import pytest

import hbseed
import oracle_lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    pq = hbseed.random_protein(3001, 800)
    res, offs = hbseed.make_db(3000, 6000, plant_query=pq, plant_frac=0.1)
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    specs = []
    lens = hbseed.lognormal_lengths(3002, 24, 350.0, 0.7, 50, 3000)
    for i, L in enumerate(lens):
        q = hbseed.random_protein(3100 + i, int(L))
        if i % 5 == 0:   # share a planted segment so that there are real homologs
            q[:min(200, len(q))] = pq[100:100 + min(200, len(q))]
        specs.append(hbseed.blastp_query(q))
    specs.append(hbseed.blastp_query(hbseed.random_protein(3200, 5000)))   # longer than a group
    for i in range(3):
        seg = hbseed.homolog(3300 + i, pq[200 + 50 * i: 300 + 50 * i], 0.7, 1)
        specs.append(hbseed.blastx_query(hbseed.blastx_nt(3310 + i, 900, seg, seg[::-1].copy())))
    yield ctx, db, res, offs, specs
    db.close()
    ctx.close()


def oracle(spec, res, offs, counters=False):
    c, h = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=8)
    h = h.view(np.int32).reshape(-1, 8)
    return (h, c) if counters else h


@pytest.mark.parametrize("group_residues", [1024, 2048])
def test_batch_equals_alone(setup, group_residues):
    ctx, db, res, offs, specs = setup
    b = hbseed.Batch(ctx, db, group_residues=group_residues)
    ids = [b.add(s) for s in specs]
    b.run()
    info = b.info()
    assert info["n_groups"] < len(specs) and info["n_classes"] >= 2
    total = hits = ext = 0
    first = {}
    for qid, spec in zip(ids, specs):
        got = b.results(qid).view(np.int32).reshape(-1, 8)
        exp, cnt = oracle(spec, res, offs, counters=True)
        assert np.array_equal(got, exp), (qid, len(got), len(exp))
        first[qid] = got.copy()
        total += len(got)
        hits += int(cnt[:, 0].sum(dtype=np.uint64))
        ext += int(cnt[:, 1].sum(dtype=np.uint64))
    assert total == info["hsps"] and total > 0
    # the launch's totals (summed over the wavefronts' per-lane counters)
    assert info["lookup_hits"] == hits and info["extensions"] == ext
    b.run()   # graph replay
    for qid in ids:
        assert np.array_equal(b.results(qid).view(np.int32).reshape(-1, 8), first[qid])
    b.close()


def test_batch_autotune(setup):
    ctx, db, res, offs, specs = setup
    b = hbseed.Batch(ctx, db, group_residues=0, tune_tiles=40)
    ids = [b.add(s) for s in specs[:10]]
    b.run()
    assert b.info()["group_residues"] in (512, 768, 1024, 1536, 2048, 3072, 4096)
    for qid, spec in zip(ids, specs[:10]):
        assert np.array_equal(b.results(qid).view(np.int32).reshape(-1, 8), oracle(spec, res, offs))
    b.close()


def test_batch_and_serve_with_long_subjects():
    """A database with subjects up to 20,000 residues: positions no longer fit
    the 16-bit diagonal state, so the batched and served launches keep the
    32-bit state; every query still equals the oracle alone."""
    pq = hbseed.random_protein(3701, 600)
    res, offs = hbseed.make_db(3700, 3000, plant_query=pq, plant_frac=0.1)
    extra = [hbseed.random_protein(3790 + i, n) for i, n in enumerate((9000, 15000, 20000))]
    extra[1][7000:7600] = pq                          # a homolog inside a long subject
    res = np.concatenate([res] + extra)
    offs = np.concatenate([offs, offs[-1] + np.cumsum([len(e) for e in extra]).astype(np.uint64)])
    assert int((offs[1:] - offs[:-1]).max()) == 20000
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    try:
        lens = hbseed.lognormal_lengths(3702, 12, 350.0, 0.7, 50, 3000)
        specs = [hbseed.blastp_query(hbseed.random_protein(3710 + i, int(L)))
                 for i, L in enumerate(lens)]
        specs[0] = hbseed.blastp_query(pq)
        b = hbseed.Batch(ctx, db, group_residues=2048)
        ids = [b.add(s) for s in specs]
        b.run()
        for qid, spec in zip(ids, specs):
            assert np.array_equal(b.results(qid).view(np.int32).reshape(-1, 8),
                                  oracle(spec, res, offs)), qid
        b.close()
        with hbseed.Server(ctx, db) as srv:
            tickets = [srv.submit(s) for s in specs[:4]]
            for t, spec in zip(tickets, specs[:4]):
                assert srv.wait(t, 120000)
                assert np.array_equal(srv.results(t).view(np.int32).reshape(-1, 8),
                                      oracle(spec, res, offs))
        q = hbseed.Query(ctx, specs[0])
        s = hbseed.Scan(ctx, db, q)
        assert s.state_mode == hbseed.STATE_LDS   # 32-bit entries for these subjects
        s.close()
        q.close()
    finally:
        db.close()
        ctx.close()


@pytest.mark.parametrize("waves,ring,items", [(1, 256, 1), (2, 1024, 16), (4, 2048, 256)])
def test_launch_geometry_keeps_results(setup, waves, ring, items):
    """The knobs the autotuner sets (wavefronts per workgroup, staging ring,
    work items per workgroup) change only speed: every query equals the
    oracle under each geometry, and the autotuned run reports its choice."""
    ctx, db, res, offs, specs = setup
    b = hbseed.Batch(ctx, db, group_residues=1024, waves_per_workgroup=waves, ring_bytes=ring,
                     items_per_workgroup=items)
    ids = [b.add(s) for s in specs]
    b.run()
    info = b.info()
    assert info["waves_per_workgroup"] == waves and info["items_per_workgroup"] == items
    assert 256 <= info["ring_bytes"] <= ring
    for qid, spec in zip(ids, specs):
        assert np.array_equal(b.results(qid).view(np.int32).reshape(-1, 8),
                              oracle(spec, res, offs)), (waves, ring, items, qid)
    b.close()
    b = hbseed.Batch(ctx, db)   # autotuned
    for s in specs:
        b.add(s)
    b.run()
    info = b.info()
    assert info["tune_candidates"] >= 3
    assert info["waves_per_workgroup"] in (1, 2, 4) and info["items_per_workgroup"] in (16, 64, 256)
    b.close()
