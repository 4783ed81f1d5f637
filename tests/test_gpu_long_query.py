"""GPU parity of the diagonal-state placements: a scan that keeps the
two-hit state in LDS and one that keeps it in a global-memory array per
wavefront (long queries) give identical HSPs and counters, both equal to
the oracle, for queries from 300 residues to beyond what LDS can hold.
"""
import numpy as np
import pytest

import hbseed
import oracle_lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = hbseed.Context(0)
    yield c
    c.close()


def run(ctx, db, spec, mode, window=40, interval=0):
    q = hbseed.Query(ctx, spec, window=window)
    try:
        q.set_state_mode(mode)
        assert q.state_mode == mode
        if interval:
            q.set_reset_interval(interval)
        s = hbseed.Scan(ctx, db, q)
        try:
            assert s.state_mode == mode
            stats, hsps = s.results()
        finally:
            s.close()
    finally:
        q.close()
    return stats, hsps


def long_spec(kind, n, seed):
    if kind == "blastp":
        return hbseed.blastp_query(hbseed.random_protein(seed, n))
    nt = hbseed.blastx_nt(seed, n, hbseed.random_protein(seed + 1, 100),
                          hbseed.random_protein(seed + 2, 100))
    return hbseed.blastx_query(nt)


@pytest.mark.parametrize("kind,n,lds", [
    ("blastp", 300, True), ("blastp", 2000, True), ("blastx", 6000, True),
    ("blastp", 20000, True), ("blastp", 40000, False), ("blastp", 60000, False),
])
def test_state_placements_match_each_other_and_oracle(ctx, kind, n, lds):
    spec = long_spec(kind, n, 3000 + n)
    print(f"{kind} query of {len(spec.seq)} residues ({spec.n_ctx} contexts)", flush=True)
    res, offs = hbseed.make_db(3100 + n, 2000, plant_query=spec.seq[:min(len(spec.seq), 3000)],
                               plant_frac=0.2)
    db = hbseed.Database(ctx, res, offs)
    try:
        sg, hg = run(ctx, db, spec, hbseed.STATE_GLOBAL)
        for mode, fits in ((hbseed.STATE_LDS, lds),):
            if fits:
                sl, hl = run(ctx, db, spec, mode)
                assert np.array_equal(hl.view(np.int32), hg.view(np.int32)), mode
                assert np.array_equal(sl.view(np.uint8), sg.view(np.uint8)), mode
                continue
            q = hbseed.Query(ctx, spec)
            try:
                with pytest.raises(hbseed.HBError) as e:
                    q.set_state_mode(mode)
                assert e.value.code == -7
                s = hbseed.Scan(ctx, db, q)   # automatic placement
                assert s.state_mode != mode
                s.close()
            finally:
                q.close()
    finally:
        db.close()
    oq = oracle_lib.OracleQuery.from_spec(spec)
    cnt, ref = oq.scan(res, offs, threads=8)
    assert len(hg) > 0
    assert np.array_equal(hg.view(np.int32), ref.view(np.int32))
    assert np.array_equal(sg["lookup_hits"], cnt[:, 0])
    assert np.array_equal(sg["extensions"], cnt[:, 1])
    assert np.array_equal(sg["n_hsp"], cnt[:, 2])


def test_gapped_stage_on_a_long_query(ctx):
    spec = hbseed.blastp_query(hbseed.random_protein(3301, 36000))
    res, offs = hbseed.make_db(3302, 1500, plant_query=spec.seq[30000:33000], plant_frac=0.3)
    db = hbseed.Database(ctx, res, offs)
    q = hbseed.Query(ctx, spec)
    try:
        s = hbseed.Scan(ctx, db, q)
        try:
            assert s.state_mode == hbseed.STATE_GLOBAL
            _, hsps = s.results()
            gst, g = s.gapped([60])
        finally:
            s.close()
    finally:
        q.close()
        db.close()
    oq = oracle_lib.OracleQuery.from_spec(spec)
    _, ref_hsps = oq.scan(res, offs, threads=8)
    assert np.array_equal(hsps.view(np.int32), ref_hsps.view(np.int32))
    cnt, ref = oracle_lib.gapped_db(oq, res, offs, ref_hsps, [60], threads=8)
    assert len(g) > 0
    assert np.array_equal(gst["n_hsp"], cnt[:, 0])
    assert np.array_equal(g.view(np.int32), ref.view(np.int32))


def test_batched_and_served_long_queries(ctx):
    """Long queries in the batched and served forms (their size classes keep
    the diagonal state in HBM) give the plain path's HSPs, next to short
    queries that keep it in LDS."""
    pq = hbseed.random_protein(3401, 2000)
    res, offs = hbseed.make_db(3400, 3000, plant_query=pq, plant_frac=0.2)
    specs = [hbseed.blastp_query(hbseed.random_protein(3410 + i, 150 + 90 * i)) for i in range(6)]
    long_p = hbseed.random_protein(3420, 40000)
    long_p[5000:7000] = pq
    specs.append(hbseed.blastp_query(long_p))
    mid = hbseed.random_protein(3421, 20000)
    mid[:2000] = pq
    specs.append(hbseed.blastp_query(mid))
    specs.append(hbseed.blastx_query(hbseed.blastx_nt(3422, 12000, pq[:100], pq[500:600])))
    db = hbseed.Database(ctx, res, offs)
    try:
        plain = [hbseed.search(ctx, db, s)[1] for s in specs]
        assert sum(len(p) for p in plain[6:]) > 0
        b = hbseed.Batch(ctx, db, group_residues=2048)
        ids = [b.add(s) for s in specs]
        b.run()
        print("batch", b.info(), flush=True)
        for qid, p in zip(ids, plain):
            assert np.array_equal(b.results(qid).view(np.int32), p.view(np.int32)), qid
        b.run()   # graph replay
        for qid, p in zip(ids, plain):
            assert np.array_equal(b.results(qid).view(np.int32), p.view(np.int32)), qid
        b.close()
        with hbseed.Server(ctx, db) as srv:
            tickets = [srv.submit(s) for s in specs]
            for t, p in zip(tickets, plain):
                assert srv.wait(t, 120000)
                assert np.array_equal(srv.results(t).view(np.int32), p.view(np.int32))
    finally:
        db.close()


def test_state_resets_keep_results(ctx):
    """The scan clears its diagonal state whenever the positions it stores
    (subject offsets accumulated over a wavefront's subjects) would pass the
    placement's limit, which the synthetic databases never reach.  Reset
    intervals down to the smallest allowed one (twice the longest subject
    plus the window) force a reset every few subjects: subjects up to 7,880
    residues, many near it, windows 40 and 255, LDS and global placements
    all give the oracle's results; a smaller interval is refused."""
    rng = np.random.default_rng(3500)
    lens = np.concatenate([rng.integers(7800, 7881, 40), rng.integers(30, 600, 400)])
    rng.shuffle(lens)
    offs = np.zeros(len(lens) + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    q = hbseed.random_protein(3501, 3000)
    res = np.concatenate([hbseed.random_protein(3502 + i, int(n)) for i, n in enumerate(lens)])
    for i in range(0, len(lens), 7):   # planted copies of query segments
        a, n = int(offs[i]), int(lens[i])
        k = min(n, 200)
        res[a + n - k:a + n] = q[100:100 + k]
    spec = hbseed.blastp_query(q)
    longest = int(lens.max())
    db = hbseed.Database(ctx, res, offs)
    try:
        for window in (40, 255):
            _, ref = oracle_lib.OracleQuery.from_spec(spec, window=window).scan(res, offs,
                                                                                threads=8)
            assert len(ref) > 0
            s0, h0 = run(ctx, db, spec, hbseed.STATE_LDS, window)
            assert np.array_equal(h0.view(np.int32), ref.view(np.int32)), window
            smallest = 2 * (longest + window) + 65
            for mode in (hbseed.STATE_LDS, hbseed.STATE_GLOBAL):
                for interval in (smallest, smallest + 5000, 60000):
                    print(f"window {window} mode {mode} reset interval {interval}", flush=True)
                    si, hi = run(ctx, db, spec, mode, window, interval)
                    assert np.array_equal(hi.view(np.int32), ref.view(np.int32)), interval
                    assert np.array_equal(si.view(np.uint8), s0.view(np.uint8)), interval
            qq = hbseed.Query(ctx, spec, window=window)
            try:
                qq.set_reset_interval(smallest - 1)
                with pytest.raises(hbseed.HBError) as e:
                    hbseed.Scan(ctx, db, qq)
                assert e.value.code == -7
            finally:
                qq.close()
    finally:
        db.close()


def test_lds16_request_is_the_lds_placement(ctx):
    """The 16-bit request of the earlier interface selects the one LDS format."""
    spec = hbseed.blastp_query(hbseed.random_protein(3601, 400))
    res, offs = hbseed.make_db(3602, 500, plant_query=spec.seq, plant_frac=0.2)
    db = hbseed.Database(ctx, res, offs)
    q = hbseed.Query(ctx, spec)
    try:
        q.set_state_mode(hbseed.STATE_LDS16)
        assert q.state_mode == hbseed.STATE_LDS
        s = hbseed.Scan(ctx, db, q)
        try:
            assert s.state_mode == hbseed.STATE_LDS
            _, h = s.results()
        finally:
            s.close()
    finally:
        q.close()
        db.close()
    _, ref = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=8)
    assert np.array_equal(h.view(np.int32), ref.view(np.int32))
