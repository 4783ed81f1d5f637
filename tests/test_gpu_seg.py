"""GPU parity of SEG-masked queries: a query with low-complexity stretches,
masked by hb_seg_mask (hard masking: lookup-table mask plus X residues, as
BLAST+ applies -seg), scans bit-exactly like the oracle scanning the same
masked query, for BLASTP and six-frame BLASTX, in the plain and the served
forms; the mask removes the low-complexity hits and keeps the homolog's.
"""
import numpy as np
import pytest

import hbseed
import oracle_lib

pytestmark = pytest.mark.gpu


def scan_both(ctx, db, spec, res, offs):
    stats, hsps = hbseed.search(ctx, db, spec)
    cnt, ref = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=8)
    assert np.array_equal(hsps.view(np.int32), ref.view(np.int32))
    assert np.array_equal(stats["lookup_hits"], cnt[:, 0])
    assert np.array_equal(stats["extensions"], cnt[:, 1])
    assert np.array_equal(stats["n_hsp"], cnt[:, 2])
    return hsps


def test_seg_masked_queries_match_oracle():
    core = hbseed.random_protein(5101, 400)
    polyq = np.full(40, 15, np.uint8)                         # Q
    dip = np.array([17, 7] * 15, np.uint8)                    # SG repeat
    q = np.concatenate([core[:150], polyq, core[150:300], dip, core[300:]])
    res, offs = hbseed.make_db(5100, 20000, plant_query=core, plant_frac=0.05)
    # low-complexity subjects that an unmasked query hits everywhere
    lc = [np.concatenate([hbseed.random_protein(5200 + i, 100), np.full(60, 15, np.uint8),
                          np.array([17, 7] * 20, np.uint8), hbseed.random_protein(5300 + i, 100)])
          for i in range(200)]
    res = np.concatenate([res] + lc)
    offs = np.concatenate([offs, offs[-1] + np.cumsum([x.size for x in lc]).astype(np.uint64)])
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    try:
        plain = hbseed.blastp_query(q)
        masked = hbseed.seg_query(plain)
        m = masked.lut_mask
        print(f"blastp: {int(m.sum())} of {q.size} residues masked", flush=True)
        assert m[150:190].all() and m[340:370].all()
        h0 = scan_both(ctx, db, plain, res, offs)
        h1 = scan_both(ctx, db, masked, res, offs)
        n_lc0 = int((h0["oid"] >= 20000).sum())
        n_lc1 = int((h1["oid"] >= 20000).sum())
        print(f"HSPs: unmasked {h0.size} ({n_lc0} on low-complexity subjects), masked {h1.size} "
              f"({n_lc1})", flush=True)
        assert n_lc1 < n_lc0 and h1.size > 100
        # six-frame BLASTX with a CAG repeat (poly-Q in one frame)
        rng = np.random.default_rng(5102)
        nt = "".join(rng.choice(list("ACGT"), 500)) + "CAG" * 50 + \
            "".join(rng.choice(list("ACGT"), 500))
        xs = hbseed.blastx_query(nt)
        xm = hbseed.seg_query(xs)
        print(f"blastx: {int(xm.lut_mask.sum())} residues masked", flush=True)
        assert xm.lut_mask.sum() >= 3 * 40
        x0 = scan_both(ctx, db, xs, res, offs)
        x1 = scan_both(ctx, db, xm, res, offs)
        assert int((x1["oid"] >= 20000).sum()) < int((x0["oid"] >= 20000).sum())
        # the same masked queries served together with unmasked ones: each
        # result equals its plain scan
        with hbseed.Server(ctx, db) as srv:
            specs = [masked, xm, plain, xs]
            tickets = [srv.submit(sp) for sp in specs]
            for t, want in zip(tickets, [h1, x1, h0, x0]):
                assert srv.wait(t, 120000)
                assert np.array_equal(srv.results(t).view(np.int32), want.view(np.int32))
    finally:
        db.close()
        ctx.close()
