"""GPU parity of the preliminary gapped extension (hb_scan_gapped) against
the oracle's restatement of BLAST_GetGappedScore, bit-exact.

The oracle extends the oracle's own ungapped HSPs (which equal the device's,
checked here again); the device extends the HSPs its scan left on the
device.  Compared exactly: per-subject (saved, extensions) and every gapped
record (oid, context, q/s ranges, start point, score) in (oid, save) order.
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


def run_both(ctx, spec, res, offs, cutoffs, gap_open=11, gap_extend=1, x_dropoff=38,
             threads=8):
    db = hbseed.Database(ctx, res, offs)
    q = hbseed.Query(ctx, spec)
    try:
        s = hbseed.Scan(ctx, db, q)
        try:
            _, hsps = s.results()
            gst, g = s.gapped(cutoffs, gap_open, gap_extend, x_dropoff)
            ms, n_wide = s.gapped_info()
        finally:
            s.close()
    finally:
        q.close()
        db.close()
    oq = oracle_lib.OracleQuery.from_spec(spec)
    _, ref_hsps = oq.scan(res, offs, threads=threads)
    assert np.array_equal(hsps.view(np.int32), ref_hsps.view(np.int32))
    cnt, ref = oracle_lib.gapped_db(oq, res, offs, ref_hsps, cutoffs, gap_open, gap_extend,
                                    x_dropoff, threads=threads)
    assert np.array_equal(gst["n_hsp"], cnt[:, 0])
    assert np.array_equal(gst["extensions"], cnt[:, 1])
    assert len(g) == len(ref), (len(g), len(ref))
    got = g.view(np.int32).reshape(-1, 10)
    exp = ref.view(np.int32).reshape(-1, 10)
    bad = np.nonzero((got != exp).any(axis=1))[0]
    assert bad.size == 0, f"first mismatch at {bad[:5]}: got {g[bad[:3]]} exp {ref[bad[:3]]}"
    first = np.full(len(offs) - 1, -1, np.int64)
    if len(ref):
        starts = np.nonzero(np.r_[True, ref["oid"][1:] != ref["oid"][:-1]])[0]
        first[ref["oid"][starts]] = starts
    assert np.array_equal(gst["first"], first)
    return gst, g, ms, n_wide


@pytest.mark.parametrize("qlen,n_subj,seed,cut", [(300, 20000, 1000, 60), (1000, 5000, 1100, 80),
                                                   (64, 3000, 7, 45), (2000, 2000, 9, 90)])
def test_gapped_blastp_matches_oracle(ctx, qlen, n_subj, seed, cut):
    q = hbseed.random_protein(seed + 1, qlen)
    res, offs = hbseed.make_db(seed, n_subj, plant_query=q, plant_frac=0.1)
    gst, g, ms, n_wide = run_both(ctx, hbseed.blastp_query(q), res, offs, [cut])
    assert len(g) > 0 and gst["extensions"].sum() > len(g)


def test_gapped_blastx_matches_oracle(ctx):
    pq = hbseed.random_protein(1201, 1000)
    res, offs = hbseed.make_db(1200, 4000, plant_query=pq, plant_frac=0.1)
    nt = hbseed.blastx_nt(1202, 1500, pq[100:200], pq[500:600])
    spec = hbseed.blastx_query(nt)
    cut = [50, 55, 60, 50, 55, 60]
    gst, g, _, _ = run_both(ctx, spec, res, offs, cut)
    assert len(g) > 0 and len(set(g["context"].tolist())) > 1


@pytest.mark.parametrize("go,ge,x", [(10, 2, 20), (9, 1, 60), (11, 1, 5), (5, 2, 0)])
def test_gapped_other_parameters(ctx, go, ge, x):
    q = hbseed.random_protein(3301, 500)
    res, offs = hbseed.make_db(3300, 3000, plant_query=q, plant_frac=0.15)
    run_both(ctx, hbseed.blastp_query(q), res, offs, [50], go, ge, x)


def test_gapped_wide_pass_matches_oracle(ctx):
    """A drop-off so large that the band outgrows the LDS ring: every
    extended subject takes the wide (global-memory) pass."""
    q = hbseed.random_protein(3401, 200)
    res, offs = hbseed.make_db(3400, 300, median=200, hi=700, plant_query=q, plant_frac=0.3)
    gst, g, _, n_wide = run_both(ctx, hbseed.blastp_query(q), res, offs, [60], x_dropoff=600)
    assert n_wide > 0 and len(g) > 0


def test_gapped_subject_with_many_hsps(ctx):
    """More ungapped HSPs in one subject than the LDS order holds (256) and
    more saved gapped HSPs than the LDS list holds (64)."""
    q = hbseed.random_protein(7101, 400)
    parts = []
    for k in range(420):
        a = (37 * k) % 340
        parts.append(hbseed.homolog(7200 + k, q[a:a + 60], 0.9, 0))
        parts.append(hbseed.random_protein(7300 + k, 7))
    big = np.concatenate(parts).astype(np.uint8)
    small = hbseed.random_protein(7400, 300)
    mid = np.concatenate(parts[:200]).astype(np.uint8)
    res = np.concatenate([small, big, mid, small]).astype(np.uint8)
    lens = [len(small), len(big), len(mid), len(small)]
    offs = np.zeros(len(lens) + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    gst, g, _, n_wide = run_both(ctx, hbseed.blastp_query(q), res, offs, [60])
    assert n_wide >= 1 and gst["n_hsp"][1] > 64


def test_gapped_long_subject(ctx):
    """A subject of more than 90000 residues (the AdjustSubjectRange window)."""
    q = hbseed.random_protein(5501, 600)
    core = hbseed.homolog(5502, q[100:400], 0.7, 2)
    pad = hbseed.random_protein(5503, 120000)
    longs = np.concatenate([pad[:95000], core, pad[95000:]]).astype(np.uint8)
    res2, offs2 = hbseed.make_db(5504, 200, plant_query=q, plant_frac=0.2)
    res = np.concatenate([longs, res2]).astype(np.uint8)
    offs = np.concatenate([[0], len(longs) + offs2.astype(np.int64)]).astype(np.uint64)
    gst, g, _, _ = run_both(ctx, hbseed.blastp_query(q), res, offs, [60])
    assert gst["n_hsp"][0] >= 1


def test_gapped_no_hsps(ctx):
    q = hbseed.random_protein(6601, 20)
    res, offs = hbseed.make_db(6600, 50)
    gst, g, _, _ = run_both(ctx, hbseed.blastp_query(q), res, offs, [60])
    assert len(g) == 0 and np.all(gst["first"] == -1)


def test_gapped_search_on_held_hsps_equals_scan_gapped(ctx):
    """hb_gapped_search on the scan's HSPs (shuffled) equals hb_scan_gapped."""
    q = hbseed.random_protein(8801, 700)
    res, offs = hbseed.make_db(8800, 4000, plant_query=q, plant_frac=0.15)
    spec = hbseed.blastp_query(q)
    db = hbseed.Database(ctx, res, offs)
    qq = hbseed.Query(ctx, spec)
    try:
        s = hbseed.Scan(ctx, db, qq)
        _, hsps = s.results()
        gst, g = s.gapped([60])
        s.close()
        perm = np.random.default_rng(1).permutation(len(hsps))
        gst2, g2 = hbseed.gapped_search(ctx, db, spec, hsps[perm], [60])
    finally:
        qq.close()
        db.close()
    assert len(g) > 0
    assert np.array_equal(g.view(np.int32), g2.view(np.int32))
    assert np.array_equal(gst.view(np.int64), gst2.view(np.int64))


def test_gapped_search_on_batched_results(ctx):
    """The gapped stage on each query's results from the batched form equals
    the oracle's gapped stage on that query."""
    pq = hbseed.random_protein(8901, 900)
    res, offs = hbseed.make_db(8900, 3000, plant_query=pq, plant_frac=0.15)
    specs = [hbseed.blastp_query(pq[a:a + L]) for a, L in [(0, 300), (200, 500), (600, 250)]]
    specs.append(hbseed.blastx_query(hbseed.blastx_nt(8902, 900, pq[100:180], pq[400:480])))
    db = hbseed.Database(ctx, res, offs)
    try:
        b = hbseed.Batch(ctx, db, group_residues=1024)
        ids = [b.add(sp) for sp in specs]
        b.run()
        got = []
        for sp, qid in zip(specs, ids):
            h = b.results(qid)
            cut = np.full(sp.n_ctx, 45, np.int32)
            got.append((h, cut) + hbseed.gapped_search(ctx, db, sp, h, cut))
        b.close()
    finally:
        db.close()
    for sp, (h, cut, gst, g) in zip(specs, got):
        oq = oracle_lib.OracleQuery.from_spec(sp)
        _, ref_h = oq.scan(res, offs, threads=8)
        assert np.array_equal(h.view(np.int32), ref_h.view(np.int32))
        cnt, ref = oracle_lib.gapped_db(oq, res, offs, ref_h, cut, threads=8)
        assert np.array_equal(g.view(np.int32), ref.view(np.int32))
        assert np.array_equal(gst["extensions"], cnt[:, 1])
        assert len(ref) > 0


def test_gapped_search_rejects_bad_records(ctx):
    q = hbseed.random_protein(9001, 100)
    res, offs = hbseed.make_db(9000, 20)
    spec = hbseed.blastp_query(q)
    db = hbseed.Database(ctx, res, offs)
    try:
        good = np.zeros(1, hbseed.HSP_DTYPE)
        good[0] = (3, 0, 10, 5, 12, 7, 8, 30)
        hbseed.gapped_search(ctx, db, spec, good, [40])
        for field, value in [("oid", 20), ("context", 1), ("q_start", 95), ("s_off", 9),
                             ("len", 0), ("s_start", -1)]:
            bad = good.copy()
            bad[field] = value
            with pytest.raises(hbseed.HBError) as e:
                hbseed.gapped_search(ctx, db, spec, bad, [40])
            assert e.value.code == -1, field
    finally:
        db.close()
