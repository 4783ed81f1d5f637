"""GPU parity: the HIP stage against the CPU oracle, bit-exact.

Every comparison is exact: HSP records (oid, context, q_start, s_start,
q_off, s_off, len, score) in (oid, s_off, q_off) order, and the per-subject
counters (lookup hits, extensions, HSPs).
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


def compare(ctx, spec, res, offs, threshold=11, window=40, matrix=None, threads=8):
    db = hbseed.Database(ctx, res, offs)
    try:
        stats, hsps = hbseed.search(ctx, db, spec, threshold=threshold, window=window,
                                    matrix=matrix)
    finally:
        db.close()
    oq = oracle_lib.OracleQuery.from_spec(spec, threshold=threshold, window=window,
                                          matrix=matrix)
    cnt, ref = oq.scan(res, offs, threads=threads)
    assert len(hsps) == len(ref), (len(hsps), len(ref))
    got = hsps.view(np.int32).reshape(-1, 8)
    exp = ref.view(np.int32).reshape(-1, 8)
    bad = np.nonzero((got != exp).any(axis=1))[0]
    assert bad.size == 0, f"first mismatch at {bad[:5]}: got {hsps[bad[:3]]} exp {ref[bad[:3]]}"
    assert np.array_equal(stats["lookup_hits"], cnt[:, 0])
    assert np.array_equal(stats["extensions"], cnt[:, 1])
    assert np.array_equal(stats["n_hsp"], cnt[:, 2])
    first = np.full(len(offs) - 1, -1, np.int64)
    if len(ref):
        starts = np.nonzero(np.r_[True, ref["oid"][1:] != ref["oid"][:-1]])[0]
        first[ref["oid"][starts]] = starts
    assert np.array_equal(stats["first"], first)
    return stats, hsps


@pytest.mark.parametrize("qlen,n_subj,seed", [(300, 20000, 1000), (1000, 5000, 1100),
                                               (64, 3000, 7), (2000, 2000, 9)])
def test_blastp_matches_oracle(ctx, qlen, n_subj, seed):
    q = hbseed.random_protein(seed + 1, qlen)
    res, offs = hbseed.make_db(seed, n_subj, plant_query=q, plant_frac=0.1)
    stats, hsps = compare(ctx, hbseed.blastp_query(q), res, offs)
    assert len(hsps) > 0


def test_blastx_matches_oracle(ctx):
    pq = hbseed.random_protein(1201, 1000)
    res, offs = hbseed.make_db(1200, 4000, plant_query=pq, plant_frac=0.1)
    # two coding regions: homologs of two 100-aa database segments
    lens = np.diff(offs)
    big = np.nonzero(lens >= 150)[0]
    s1 = res[offs[big[0]]:offs[big[0]] + 100]
    s2 = res[offs[big[1]] + 20:offs[big[1]] + 120]
    nt = hbseed.blastx_nt(1202, 1500, hbseed.homolog(1203, s1, 0.7, 1),
                          hbseed.homolog(1204, s2, 0.7, 1))
    spec = hbseed.blastx_query(nt)
    stats, hsps = compare(ctx, spec, res, offs)
    assert set(np.unique(hsps["context"])) >= {0, 3} or len(hsps) > 0


def test_edge_subjects(ctx):
    q = hbseed.random_protein(5, 200)
    parts = [np.zeros(0, np.uint8), q[:1], q[:2], q[:3], q[10:80], q[:200],
             np.full(500, 1, np.uint8), q[::-1].copy(), hbseed.random_protein(6, 4000)]
    res = np.concatenate(parts)
    offs = np.zeros(len(parts) + 1, np.uint64)
    offs[1:] = np.cumsum([len(p) for p in parts])
    compare(ctx, hbseed.blastp_query(q), res, offs)


def test_repetitive_query(ctx):
    # long runs make large lookup cells and many hits on one diagonal
    q = np.concatenate([np.full(150, 1, np.uint8), hbseed.random_protein(8, 100),
                        np.tile(np.array([9, 11, 19], np.uint8), 50)])
    res, offs = hbseed.make_db(11, 2000, plant_query=q, plant_frac=0.2)
    extra = np.concatenate([np.full(300, 1, np.uint8), np.tile(np.array([9, 11, 19], np.uint8), 100)])
    res = np.concatenate([res, extra])
    offs = np.append(offs, offs[-1] + len(extra)).astype(np.uint64)
    compare(ctx, hbseed.blastp_query(q), res, offs)


def test_device_lut_matches_oracle(ctx):
    q = hbseed.random_protein(21, 700)
    q[5:8] = [21, 25, 21]   # X * X: self-score below T, must still seed itself
    spec = hbseed.blastp_query(q)
    dq = hbseed.Query(ctx, spec)
    try:
        start, ent = dq.lut()
    finally:
        dq.close()
    oq = oracle_lib.OracleQuery.from_spec(spec)
    ostart, oent = oq.lut_arrays()
    assert np.array_equal(start, ostart)
    assert np.array_equal(ent, oent)


def masked_spec(spec, seed, runs=6, longest=15):
    rng = np.random.default_rng(seed)
    mask = np.zeros(len(spec.seq), np.uint8)
    for a in rng.integers(0, len(spec.seq) - 3, runs):
        mask[a:a + int(rng.integers(3, longest))] = 1
    return hbseed.QuerySpec(spec.seq, spec.ctx_off, spec.ctx_len, spec.xdrop, spec.cutoff,
                            lut_mask=mask)


def test_masked_lut_and_scan_match_oracle(ctx):
    """A lookup-table mask (NCBI's masked segments) on the device equals the
    oracle: table entries, HSPs and counters."""
    q = hbseed.random_protein(4101, 600)
    spec = masked_spec(hbseed.blastp_query(q), 4102, runs=12, longest=40)
    dq = hbseed.Query(ctx, spec)
    try:
        start, ent = dq.lut()
    finally:
        dq.close()
    ostart, oent = oracle_lib.OracleQuery.from_spec(spec).lut_arrays()
    assert np.array_equal(start, ostart) and np.array_equal(ent, oent)
    res, offs = hbseed.make_db(4100, 5000, plant_query=q, plant_frac=0.1)
    stats, hsps = compare(ctx, spec, res, offs)
    _, unmasked = compare(ctx, hbseed.blastp_query(q), res, offs)
    assert 0 < len(hsps) < len(unmasked)


def test_masked_blastx_and_batch_match_oracle(ctx):
    pq = hbseed.random_protein(4201, 800)
    res, offs = hbseed.make_db(4200, 3000, plant_query=pq, plant_frac=0.1)
    specs = [masked_spec(hbseed.blastx_query(hbseed.blastx_nt(4202, 1200, pq[100:200],
                                                              pq[400:500])), 4203),
             masked_spec(hbseed.blastp_query(pq[:500]), 4204), hbseed.blastp_query(pq[300:700])]
    compare(ctx, specs[0], res, offs)
    db = hbseed.Database(ctx, res, offs)
    try:
        b = hbseed.Batch(ctx, db, group_residues=1024)
        ids = [b.add(sp) for sp in specs]
        b.run()
        got = [b.results(i) for i in ids]
        b.close()
    finally:
        db.close()
    for sp, h in zip(specs, got):
        _, ref = oracle_lib.OracleQuery.from_spec(sp).scan(res, offs, threads=8)
        assert np.array_equal(h.view(np.int32), ref.view(np.int32))


def test_long_subjects_global_path(ctx):
    # subjects longer than the LDS staging buffer take the global-memory path
    q = hbseed.random_protein(31, 400)
    parts = [hbseed.random_protein(32, 6000), hbseed.random_protein(33, 10000)]
    for i, seed in enumerate((34, 35)):
        s = hbseed.random_protein(seed, 9000)
        h = hbseed.homolog(seed + 100, q[50:170], 0.6, 2)
        s[4500 + i:4500 + i + len(h)] = h
        parts.append(s)
    res = np.concatenate(parts)
    offs = np.zeros(len(parts) + 1, np.uint64)
    offs[1:] = np.cumsum([len(p) for p in parts])
    stats, hsps = compare(ctx, hbseed.blastp_query(q), res, offs)
    assert len(hsps) > 0


@pytest.mark.parametrize("nt_len", [600, 3000])
def test_blastx_lengths(ctx, nt_len):
    pq = hbseed.random_protein(41, 800)
    res, offs = hbseed.make_db(40, 3000, plant_query=pq, plant_frac=0.1)
    seg = hbseed.homolog(42, res[offs[5]:offs[5] + 60], 0.7, 0)
    nt = hbseed.blastx_nt(43 + nt_len, nt_len, seg, hbseed.homolog(44, pq[100:160], 0.6, 0))
    compare(ctx, hbseed.blastx_query(nt), res, offs)


def test_long_concatenated_query(ctx):
    # 12,000-residue query: diagonal arrays force one wavefront per workgroup
    q = hbseed.random_protein(51, 12000)
    res, offs = hbseed.make_db(50, 800, plant_query=q, plant_frac=0.2)
    compare(ctx, hbseed.blastp_query(q), res, offs)


def test_subject_with_many_hsps(ctx):
    """Subjects holding hundreds of homologous copies of query segments:
    more than 256 HSPs in one subject take the host ordering path of
    hb_scan_results; the records must still come out in (s_off, q_off)
    order and equal the oracle."""
    q = hbseed.random_protein(7101, 400)
    parts = []
    for k in range(420):
        a = (37 * k) % 340
        seg = hbseed.homolog(7200 + k, q[a:a + 60], 0.9, 0)
        parts.append(seg)
        parts.append(hbseed.random_protein(7300 + k, 7))
    big = np.concatenate(parts).astype(np.uint8)
    small = hbseed.random_protein(7400, 300)
    mid = np.concatenate(parts[:200]).astype(np.uint8)
    res = np.concatenate([small, big, mid, small]).astype(np.uint8)
    lens = [len(small), len(big), len(mid), len(small)]
    offs = np.zeros(len(lens) + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    stats, hsps = compare(ctx, hbseed.blastp_query(q), res, offs)
    assert stats["n_hsp"][1] > 256 and stats["n_hsp"][2] > 0


def test_plain_output_overflow_rerun(ctx):
    """More HSPs than the plain path's first reservation (2^18 records for a
    small database): the pass is rerun with the counted size and nothing is
    dropped (H-BLAST instead flags the subject for a CPU recompute,
    h_blast_kernel.cu:468-471)."""
    q = hbseed.random_protein(6101, 400)
    segs = [np.concatenate([hbseed.homolog(6200 + k, q[(29 * k) % 330:(29 * k) % 330 + 60], 0.9, 0),
                            hbseed.random_protein(6300 + k, 9)]) for k in range(97)]
    rng = np.random.default_rng(6102)
    parts = [np.concatenate([segs[j] for j in rng.integers(0, 97, 24)]) for _ in range(14000)]
    res = np.concatenate(parts).astype(np.uint8)
    offs = np.zeros(len(parts) + 1, np.uint64)
    np.cumsum([len(p) for p in parts], out=offs[1:])
    stats, hsps = compare(ctx, hbseed.blastp_query(q), res, offs, threads=8)
    assert len(hsps) > (1 << 18)


def perturbed_blosum(seed):
    """BLOSUM62 with a random symmetric +-1 perturbation on the 27 letters
    (the gap letter keeps -128, so sentinels still stop extensions)."""
    m = oracle_lib.blosum62().astype(np.int16)
    rng = np.random.default_rng(seed)
    d = rng.integers(-1, 2, size=(28, 28))
    d = np.triu(d) + np.triu(d, 1).T
    d[0, :] = 0
    d[:, 0] = 0
    return np.clip(m + d, -128, 127).astype(np.int8)


@pytest.mark.parametrize("threshold,window,matrix_seed", [(13, 40, None), (9, 40, None),
                                                          (11, 16, None), (11, 120, None),
                                                          (11, 255, None), (12, 60, 77)])
def test_search_parameters_match_oracle(ctx, threshold, window, matrix_seed):
    """Non-default neighbourhood threshold, two-hit window and score matrix
    (each reaches the device: lookup build, diagonal sizing, two-hit rule,
    extensions)."""
    q = hbseed.random_protein(9101, 500)
    res, offs = hbseed.make_db(9100, 3000, plant_query=q, plant_frac=0.15)
    m = None if matrix_seed is None else perturbed_blosum(matrix_seed)
    stats, hsps = compare(ctx, hbseed.blastp_query(q), res, offs, threshold=threshold,
                          window=window, matrix=m)
    assert len(hsps) > 0


def test_per_context_xdrop_and_cutoff(ctx):
    """BLASTX frames with their own X-drop and cutoff values."""
    pq = hbseed.random_protein(9201, 900)
    res, offs = hbseed.make_db(9200, 3000, plant_query=pq, plant_frac=0.15)
    base = hbseed.blastx_query(hbseed.blastx_nt(9202, 1500, pq[100:200], pq[500:600]))
    spec = hbseed.QuerySpec(base.seq, base.ctx_off, base.ctx_len,
                            xdrop=[10, 16, 22, 30, 7, 16], cutoff=[35, 41, 50, 30, 41, 60])
    stats, hsps = compare(ctx, spec, res, offs)
    assert len(hsps) > 0


def test_smallest_left_walk_crosses_offset_zero(ctx):
    """The two-hit case at the smallest offsets (the trigger 3 residues after
    the first hit on both sequences, no improving prefix, so the left walk
    starts at offset 2 and reaches offset 0), alone and repeated over
    subjects, equals the oracle."""
    import test_oracle_pins as pins
    m, q, s = pins.smallest_left_walk_case()
    spec = hbseed.QuerySpec(q, [0], [len(q)], [16], [5])
    res = np.concatenate([s] * 50 + [q[:6], s[:7]])
    offs = np.zeros(53, np.uint64)
    offs[1:51] = np.arange(1, 51) * len(s)
    offs[51] = offs[50] + 6
    offs[52] = offs[51] + 7
    compare(ctx, spec, res, offs, matrix=m)


def _compact_equals_dense(ctx, spec, res, offs):
    db = hbseed.Database(ctx, res, offs)
    q = hbseed.Query(ctx, spec)
    try:
        s = hbseed.Scan(ctx, db, q)
        stats, hsps = s.results()
        subj, hsps_c, tot = s.results_compact()
        s.close()
    finally:
        q.close()
        db.close()
    assert np.array_equal(hsps_c.view(np.int32).reshape(-1, 8), hsps.view(np.int32).reshape(-1, 8))
    with_hsps = np.nonzero(stats["n_hsp"])[0]
    assert np.array_equal(subj["oid"], with_hsps)
    for f in ("n_hsp", "lookup_hits", "extensions"):
        assert np.array_equal(subj[f], stats[f][with_hsps]), f
    assert np.array_equal(subj["first"].astype(np.int64), stats["first"][with_hsps])
    assert tot == {"lookup_hits": int(stats["lookup_hits"].sum(dtype=np.uint64)),
                   "extensions": int(stats["extensions"].sum(dtype=np.uint64)),
                   "n_hsp": len(hsps), "n_subjects": len(with_hsps)}
    return subj, tot


def test_compact_results_equal_dense(ctx):
    """hb_scan_results_compact: the entries of the subjects with HSPs (in oid
    order) equal hb_scan_results' per-subject stats of those subjects, the
    totals equal the sums over every subject, and the records are the same;
    also with subjects of more than 256 records (host ordering)."""
    q = hbseed.random_protein(7501, 500)
    res, offs = hbseed.make_db(7500, 30000, plant_query=q, plant_frac=0.05)
    subj, tot = _compact_equals_dense(ctx, hbseed.blastp_query(q), res, offs)
    assert 0 < tot["n_subjects"] < len(offs) - 1
    q = hbseed.random_protein(7101, 400)
    parts = []
    for k in range(420):
        a = (37 * k) % 340
        parts.append(hbseed.homolog(7200 + k, q[a:a + 60], 0.9, 0))
        parts.append(hbseed.random_protein(7300 + k, 7))
    big = np.concatenate(parts).astype(np.uint8)
    small = hbseed.random_protein(7400, 300)
    res = np.concatenate([small, big, small, big]).astype(np.uint8)
    lens = [len(small), len(big), len(small), len(big)]
    offs = np.zeros(len(lens) + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    subj, tot = _compact_equals_dense(ctx, hbseed.blastp_query(q), res, offs)
    assert subj["n_hsp"].max() > 256


def test_results_into_reused_buffers(ctx):
    """Scan.results / results_compact / gapped into one ResultBuffers each
    over scans of different result sizes (a larger one first, then a smaller
    one, then a larger one again) return the same stats, entries and records
    as fresh arrays."""
    res, offs = hbseed.make_db(7600, 20000, plant_query=hbseed.random_protein(7601, 600),
                               plant_frac=0.05)
    db = hbseed.Database(ctx, res, offs)
    bufs, cbufs, gbufs = hbseed.ResultBuffers(), hbseed.ResultBuffers(), hbseed.ResultBuffers()
    try:
        for seed, qlen in ((7601, 600), (7602, 150), (7603, 700)):
            q = hbseed.Query(ctx, hbseed.blastp_query(hbseed.random_protein(seed, qlen)))
            try:
                s = hbseed.Scan(ctx, db, q)
                stats, hsps = s.results()
                stats_r, hsps_r = s.results(bufs)
                subj, hsps_c, tot = s.results_compact()
                subj_r, hsps_cr, tot_r = s.results_compact(cbufs)
                gst, gh = s.gapped([60])
                gst_r, gh_r = s.gapped([60], out=gbufs)
                s.close()
            finally:
                q.close()
            assert np.array_equal(stats_r, stats) and np.array_equal(hsps_r, hsps)
            assert np.array_equal(subj_r, subj) and np.array_equal(hsps_cr, hsps_c)
            assert tot_r == tot and len(hsps_r) == tot["n_hsp"]
            assert np.array_equal(gst_r, gst) and np.array_equal(gh_r, gh)
    finally:
        db.close()
