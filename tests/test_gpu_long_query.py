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


def run(ctx, db, spec, mode, window=40):
    q = hbseed.Query(ctx, spec, window=window)
    try:
        q.set_state_mode(mode)
        assert q.state_mode == mode
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


@pytest.mark.parametrize("kind,n,lds,lds16", [
    ("blastp", 300, True, True), ("blastp", 2000, True, True), ("blastx", 6000, True, True),
    ("blastp", 20000, True, True), ("blastp", 40000, False, True),
    ("blastp", 60000, False, False),
])
def test_state_placements_match_each_other_and_oracle(ctx, kind, n, lds, lds16):
    spec = long_spec(kind, n, 3000 + n)
    print(f"{kind} query of {len(spec.seq)} residues ({spec.n_ctx} contexts)", flush=True)
    res, offs = hbseed.make_db(3100 + n, 2000, plant_query=spec.seq[:min(len(spec.seq), 3000)],
                               plant_frac=0.2)
    db = hbseed.Database(ctx, res, offs)
    try:
        sg, hg = run(ctx, db, spec, hbseed.STATE_GLOBAL)
        for mode, fits in ((hbseed.STATE_LDS, lds), (hbseed.STATE_LDS16, lds16)):
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


def test_16bit_state_at_its_subject_length_limit(ctx):
    """Subjects up to 7,880 residues, many near it (with window 255 the
    14-bit positions then reach 16,270 before a reset), window 40 and 255:
    the 16-bit state gives the 32-bit state's and the oracle's results;
    longer subjects refuse it."""
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
    db = hbseed.Database(ctx, res, offs)
    try:
        for window in (40, 255):
            s32, h32 = run(ctx, db, spec, hbseed.STATE_LDS, window)
            s16, h16 = run(ctx, db, spec, hbseed.STATE_LDS16, window)
            assert len(h32) > 0
            assert np.array_equal(h16.view(np.int32), h32.view(np.int32)), window
            assert np.array_equal(s16.view(np.uint8), s32.view(np.uint8)), window
            _, ref = oracle_lib.OracleQuery.from_spec(spec, window=window).scan(res, offs,
                                                                                threads=8)
            assert np.array_equal(h16.view(np.int32), ref.view(np.int32)), window
    finally:
        db.close()
    # a 9,000-residue subject: 16-bit positions no longer fit
    offs2 = np.array([0, 9000, 9100], np.uint64)
    res2 = hbseed.random_protein(3600, 9100)
    db = hbseed.Database(ctx, res2, offs2)
    try:
        q = hbseed.Query(ctx, spec)
        try:
            q.set_state_mode(hbseed.STATE_LDS16)
            with pytest.raises(hbseed.HBError) as e:
                hbseed.Scan(ctx, db, q)
            assert e.value.code == -7
        finally:
            q.close()
    finally:
        db.close()
