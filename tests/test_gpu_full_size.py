"""Bit-exact parity at full size: the 1000-residue BLASTP query against the
whole 2,000,000-subject database (718.8 M residues; the benchmark's
workload), its preliminary gapped stage, and a six-frame BLASTX query of
1,500 nucleotides against the same database -- every HSP record and every
per-subject counter against the oracle run on all host threads.
"""
import os

import numpy as np
import pytest

import hbseed
import oracle_lib

pytestmark = pytest.mark.gpu

THREADS = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or
                     len(os.sched_getaffinity(0)), 64))


@pytest.fixture(scope="module")
def c1():
    q = hbseed.random_protein(1101, 1000)
    res, offs = hbseed.make_db(1100, 2_000_000, plant_query=q, plant_frac=0.05)
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    print(f"database: {len(offs) - 1} subjects, {int(offs[-1])} residues; oracle on "
          f"{THREADS} threads", flush=True)
    yield ctx, db, q, res, offs
    db.close()
    ctx.close()


def check(stats, hsps, cnt, ref):
    assert len(hsps) == len(ref), (len(hsps), len(ref))
    got = hsps.view(np.int32).reshape(-1, 8)
    exp = ref.view(np.int32).reshape(-1, 8)
    bad = np.nonzero((got != exp).any(axis=1))[0]
    assert bad.size == 0, f"first mismatch at {bad[:5]}"
    assert np.array_equal(stats["lookup_hits"], cnt[:, 0])
    assert np.array_equal(stats["extensions"], cnt[:, 1])
    assert np.array_equal(stats["n_hsp"], cnt[:, 2])


def test_c1_blastp_full_database(c1):
    ctx, db, q, res, offs = c1
    spec = hbseed.blastp_query(q)
    query = hbseed.Query(ctx, spec)
    try:
        s = hbseed.Scan(ctx, db, query)
        try:
            stats, hsps = s.results()
            gst, g = s.gapped([82])   # E = 10 over the search space (bench cutoff)
        finally:
            s.close()
    finally:
        query.close()
    print(f"device: {len(hsps)} HSPs, {len(g)} gapped HSPs", flush=True)
    oq = oracle_lib.OracleQuery.from_spec(spec)
    cnt, ref = oq.scan(res, offs, threads=THREADS)
    check(stats, hsps, cnt, ref)
    assert len(ref) > 250_000
    gcnt, gref = oracle_lib.gapped_db(oq, res, offs, ref, [82], threads=THREADS)
    assert np.array_equal(gst["n_hsp"], gcnt[:, 0])
    assert np.array_equal(gst["extensions"], gcnt[:, 1])
    assert np.array_equal(g.view(np.int32), gref.view(np.int32))


def test_blastx_1500nt_full_database(c1):
    ctx, db, q, res, offs = c1
    # two coding regions: homologs (70% identity, one indel) of 100-residue
    # segments of two database subjects, one on each strand
    segs = []
    for k, oid in enumerate((123_457, 1_500_001)):
        a, b = int(offs[oid]), int(offs[oid + 1])
        src = res[a:a + min(100, b - a)]
        segs.append(hbseed.homolog(1201 + k, src, 0.7, 1))
    nt = hbseed.blastx_nt(1200, 1500, segs[0], segs[1])
    spec = hbseed.blastx_query(nt)
    stats, hsps = hbseed.search(ctx, db, spec)
    print(f"BLASTX: {len(spec.seq)} residues in 6 frames, {len(hsps)} HSPs", flush=True)
    cnt, ref = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=THREADS)
    check(stats, hsps, cnt, ref)
    assert set(np.unique(hsps["context"])) == set(range(6))


def test_batched_c3_queries_full_database(c1):
    """The first 24 queries of the 2,048-query batched workload (log-normal
    lengths, median 350) in one batched scan of the whole database, each
    against the oracle alone; the autotuned group size."""
    ctx, db, q, res, offs = c1
    lens = hbseed.lognormal_lengths(1300, 24, 350.0, 0.7, 50, 3000)
    specs = [hbseed.blastp_query(hbseed.random_protein(1300 * 100003 + i, int(L)))
             for i, L in enumerate(lens)]
    specs[3] = hbseed.blastp_query(q[:400].copy())   # one query with planted homologs
    b = hbseed.Batch(ctx, db)
    ids = [b.add(s) for s in specs]
    b.run()
    print("batch", b.info(), flush=True)
    total = 0
    for qid, spec in zip(ids, specs):
        got = b.results(qid)
        _, ref = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=THREADS)
        assert np.array_equal(got.view(np.int32), ref.view(np.int32)), qid
        total += len(got)
    b.close()
    assert total > 10_000
