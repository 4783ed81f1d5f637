"""The batched and served configurations at their full sizes.

* Batched: all 2,048 queries of the batched workload in one batched scan
  (autotuned launch) of the 2,000,000-subject database.  Every HSP of every
  query passes the property check (extent, context, diagonal, score
  recomputed from the sequences, cutoff, order); queries spread over the
  length range (so over the launch's size classes) equal the oracle run on
  the whole database.
* Served: the served database of 20,000,000 subjects (7.2 G residues, so
  subject offsets pass 2^32 in both residue layouts) resident on the device,
  and a mixed 80/20 BLASTP/BLASTX stream submitted from a client thread.
  Every request's HSPs pass the property check; every request equals the
  oracle on a subject sample made of the first 20,000 oids, the last 20,000
  oids (their residues start past byte 2^32 in oid order) and 150,000
  subjects whose residues start past byte 2^32 in the scan's
  processing-order copy; a plain scan of one BLASTP query gives per-subject
  counters that equal the oracle's on the same sample and satisfy
  n_hsp <= extensions <= lookup hits on all 20,000,000 subjects.
"""
import os
import threading
import time

import numpy as np
import pytest

import hbseed
import oracle_lib
from hbseed import workloads

pytestmark = pytest.mark.gpu

THREADS = max(1, min(int(os.environ.get("OMP_NUM_THREADS", "0")) or
                     len(os.sched_getaffinity(0)), 64))


def say(msg):
    print(msg, flush=True)


def property_check(spec, res, offs, hsps, what):
    bad, first = oracle_lib.OracleQuery.from_spec(spec).check_hsps(res, offs, hsps)
    assert bad == 0, f"{what}: {bad} bad HSP records, first {first}: {hsps[first]}"


def subset_db(res, offs, oids):
    """Residues and offsets of the subjects oids (ascending) as a database."""
    lens = (offs[oids + 1] - offs[oids]).astype(np.int64)
    sub_offs = np.zeros(len(oids) + 1, np.uint64)
    sub_offs[1:] = np.cumsum(lens)
    idx = np.repeat(offs[oids].astype(np.int64) - sub_offs[:-1].astype(np.int64), lens) + \
        np.arange(int(sub_offs[-1]), dtype=np.int64)
    return res[idx], sub_offs


def restrict(hsps, oids):
    """HSPs of subjects oids (ascending), renumbered to their subset index."""
    keep = np.isin(hsps["oid"], oids)
    h = hsps[keep].copy()
    h["oid"] = np.searchsorted(oids, h["oid"]).astype(np.uint32)
    return h


def test_batched_all_2048_queries():
    q0, res, offs = workloads.plain_db()
    qs = workloads.batched_queries(2048)
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    try:
        b = hbseed.Batch(ctx, db)
        specs = [hbseed.blastp_query(q) for q in qs]
        ids = [b.add(s) for s in specs]
        t0 = time.perf_counter()
        b.run()
        info = b.info()
        say(f"batched: 2048 queries, {info['n_groups']} groups, {info['n_classes']} launches, "
            f"group size {info['group_residues']}, {time.perf_counter() - t0:.1f} s")
        lens = np.array([len(q) for q in qs])
        order = np.argsort(lens, kind="stable")
        pick = sorted({int(order[0]), int(order[len(order) // 4]), int(order[len(order) // 2]),
                       int(order[3 * len(order) // 4]), int(order[-1]), 7, 1000, 2047})
        total = 0
        for i, (qid, spec) in enumerate(zip(ids, specs)):
            got = b.results(qid)
            total += len(got)
            property_check(spec, res, offs, got, f"query {i}")
            if i in pick:
                _, ref = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=THREADS)
                assert np.array_equal(got.view(np.int32), ref.view(np.int32)), i
                say(f"query {i} ({lens[i]} residues): {len(got)} HSPs equal the oracle")
            if i % 256 == 255:
                say(f"checked {i + 1} queries, {total} HSPs")
        b.close()
        assert total > 1_000_000
    finally:
        db.close()
        ctx.close()


@pytest.fixture(scope="module")
def served():
    t0 = time.perf_counter()
    res, offs = workloads.served_db()
    n = len(offs) - 1
    say(f"served database: {n} subjects, {int(offs[-1])} residues "
        f"({time.perf_counter() - t0:.1f} s)")
    assert int(offs[-1]) > 2 ** 32 + 2 ** 31
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    say("database resident")
    # the scan's processing-order copy: length-descending (ties by oid),
    # each subject's span padded to 16 bytes
    lens = (offs[1:] - offs[:-1]).astype(np.int64)
    order = np.argsort(-lens, kind="stable")
    spans = (lens[order] + 15) & ~15
    start = np.concatenate([[0], np.cumsum(spans)[:-1]])
    high = order[start >= 2 ** 32]
    rng = np.random.default_rng(1401)
    sample = np.concatenate([np.arange(20000), np.arange(n - 20000, n),
                             rng.choice(high, 150000, replace=False)])
    oids = np.unique(sample).astype(np.int64)
    assert (offs[oids[-20000:]] >= 2 ** 32).all()
    sres, soffs = subset_db(res, offs, oids)
    say(f"sample: {len(oids)} subjects, {int(soffs[-1])} residues, "
        f"{int((start[np.argsort(order)][oids] >= 2 ** 32).sum())} past 2^32 in the scan copy")
    is_high = np.zeros(n, bool)
    is_high[high] = True
    yield ctx, db, res, offs, oids, sres, soffs, is_high
    db.close()
    ctx.close()


def test_served_mixed_stream_20m_subjects(served):
    ctx, db, res, offs, oids, sres, soffs, is_high = served
    gaps, specs = workloads.served_stream(40)
    arrivals = np.cumsum(gaps / 20.0)   # 20 queries/s
    assert 5 <= sum(k == "blastx" for k, _ in specs) <= 15
    srv = hbseed.Server(ctx, db, max_batch_residues=16384, max_inflight=2)
    tickets = [None] * len(specs)
    t0 = time.perf_counter()

    def client():
        for i, (_, spec) in enumerate(specs):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(spec)

    th = threading.Thread(target=client)
    th.start()
    got = []
    for i in range(len(specs)):
        while tickets[i] is None:
            time.sleep(0.001)
        assert srv.wait(tickets[i], 600000)
        got.append(srv.results(tickets[i]))
    th.join()
    st = srv.stats()
    srv.stop()
    say(f"served {len(specs)} requests in {time.perf_counter() - t0:.1f} s, "
        f"{st['batches']} batches, p50 {st['p50_ms']:.0f} ms, p99 {st['p99_ms']:.0f} ms")
    past_oid = past_scan = 0
    for i, ((kind, spec), h) in enumerate(zip(specs, got)):
        property_check(spec, res, offs, h, f"request {i} ({kind})")
        _, ref = oracle_lib.OracleQuery.from_spec(spec).scan(sres, soffs, threads=THREADS)
        mine = restrict(h, oids)
        assert np.array_equal(mine.view(np.int32), ref.view(np.int32)), (i, kind)
        past_oid += int((offs[h["oid"]] >= 2 ** 32).sum())
        past_scan += int(is_high[h["oid"]].sum())
        say(f"request {i} ({kind}, {len(spec.seq)} residues): {len(h)} HSPs; "
            f"{len(ref)} in the sample equal the oracle")
    # HSPs of subjects read through 64-bit offsets in both layouts
    say(f"HSPs in subjects past byte 2^32: {past_oid} (oid order), {past_scan} (scan copy)")
    assert past_oid > 1000 and past_scan > 1000


def test_plain_scan_counters_20m_subjects(served):
    ctx, db, res, offs, oids, sres, soffs, _ = served
    _, specs = workloads.served_stream(3)
    spec = next(s for k, s in specs if k == "blastp")
    stats, hsps = hbseed.search(ctx, db, spec)
    say(f"plain scan: {len(hsps)} HSPs, {int(stats['lookup_hits'].sum())} lookup hits")
    assert (stats["n_hsp"] <= stats["extensions"]).all()
    assert (stats["extensions"] <= stats["lookup_hits"]).all()
    property_check(spec, res, offs, hsps, "plain scan")
    cnt, ref = oracle_lib.OracleQuery.from_spec(spec).scan(sres, soffs, threads=THREADS)
    assert np.array_equal(restrict(hsps, oids).view(np.int32), ref.view(np.int32))
    assert np.array_equal(stats["lookup_hits"][oids], cnt[:, 0])
    assert np.array_equal(stats["extensions"][oids], cnt[:, 1])
    assert np.array_equal(stats["n_hsp"][oids], cnt[:, 2])
