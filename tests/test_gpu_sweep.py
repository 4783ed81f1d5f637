"""GPU parity over a seeded random sweep of small configurations.

Each case draws its own query (BLASTP or a six-frame BLASTX query, 1 to
1,500 residues or 60 to 2,400 nt), database (subject lengths from 0 up,
log-normal with a random median and spread up to 12,000 residues, planted
homologs, runs of one residue and short tandem repeats), neighbourhood threshold, two-hit window,
X-drop and cutoff, and compares the plain path with the oracle bit-exactly
(HSP records and every subject's counters).  The batched form is checked
the same way on a mix of such queries, and the gapped stage on further
random cases with random gap costs, X-drop and cutoffs.  Sizes keep the
oracle at a fraction of a second per case.
"""
import numpy as np
import pytest

import hbseed
import oracle_lib

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

QLENS = [1, 2, 3, 4, 5, 17, 63, 64, 65, 127, 300, 700, 1500]


@pytest.fixture(scope="module")
def ctx():
    c = hbseed.Context(0)
    yield c
    c.close()


def random_db(rng, seed, query):
    """Random lengths (0 allowed), planted homologs of the query, and a few
    low-complexity subjects (runs of one residue, short tandem repeats)."""
    n = int(rng.integers(300, 2500))
    median = float(rng.uniform(20, 800))
    sigma = float(rng.uniform(0.3, 1.2))
    lo = int(rng.integers(0, 4))
    plant = float(rng.uniform(0.0, 0.3)) if len(query) >= 3 else 0.0
    # up to 12,000 residues: subjects past the 4 KB staging ring are scanned
    # from global memory
    hi = int(rng.choice([3000, 6000, 12000]))
    res, offs = hbseed.make_db(seed, n, median=median, sigma=sigma, lo=lo, hi=hi,
                               plant_query=query if plant > 0 else None, plant_frac=plant)
    # the shortest subjects: empty, and shorter than a word
    extra = [hbseed.random_protein(seed + 1, k) for k in (0, 1, 2, 3)]
    for _ in range(int(rng.integers(0, 6))):
        if rng.random() < 0.5:
            extra.append(np.full(int(rng.integers(1, 400)), rng.integers(1, 25), np.uint8))
        else:
            unit = rng.integers(1, 25, size=int(rng.integers(2, 5))).astype(np.uint8)
            extra.append(np.tile(unit, int(rng.integers(1, 120))))
    res = np.concatenate([res] + extra)
    offs = np.concatenate([offs, offs[-1] + np.cumsum([len(e) for e in extra])])
    return res, offs.astype(np.uint64)


def random_case(i):
    rng = np.random.default_rng(7700 + i)
    seed = 7700 + 13 * i
    blastx = rng.random() < 0.3
    if blastx:
        pq = hbseed.random_protein(seed + 1, 600)
        nt_len = int(rng.integers(60, 2400))
        seg = min(100, nt_len // 8)   # two coding regions that fit the query
        a, b = sorted(rng.integers(0, 500, size=2))
        nt = hbseed.blastx_nt(seed + 2, nt_len, pq[a:a + seg], pq[b:b + seg])
        base = hbseed.blastx_query(nt)
        plant_src = pq
    else:
        qlen = int(rng.choice(QLENS))
        q = hbseed.random_protein(seed + 1, qlen)
        if qlen >= 30 and rng.random() < 0.3:   # a low-complexity stretch in the query
            k = int(rng.integers(5, qlen // 3))
            q[:k] = rng.integers(1, 25)
        base = hbseed.blastp_query(q)
        plant_src = q
    n_ctx = len(base.ctx_off)
    xdrop = rng.integers(5, 41, size=n_ctx).tolist()
    cutoff = rng.integers(18, 61, size=n_ctx).tolist()
    spec = hbseed.QuerySpec(base.seq, base.ctx_off, base.ctx_len, xdrop=xdrop, cutoff=cutoff)
    threshold = int(rng.integers(9, 15))
    window = int(rng.choice([4, 5, 16, 40, 41, 100, 255]))
    res, offs = random_db(rng, seed + 3, plant_src)
    return spec, res, offs, threshold, window


@pytest.mark.parametrize("i", range(120))
def test_random_configuration_matches_oracle(ctx, i):
    spec, res, offs, threshold, window = random_case(i)
    print(f"case {i}: query {len(spec.seq)} residues in {len(spec.ctx_off)} contexts, "
          f"{len(offs) - 1} subjects, T={threshold} A={window}", flush=True)
    compare(ctx, spec, res, offs, threshold=threshold, window=window)


def test_random_batch_matches_oracle(ctx):
    """A batched launch over queries of every shape the sweep draws (default
    threshold and window, each query with its own X-drop and cutoff)."""
    rng = np.random.default_rng(7800)
    res, offs = hbseed.make_db(7801, 3000, median=250, sigma=0.9, lo=0, hi=6000,
                               plant_query=hbseed.random_protein(7802, 800), plant_frac=0.2)
    specs = []
    for k in range(16):
        if k % 4 == 3:
            nt_len = int(rng.integers(60, 2000))
            seg = min(60, nt_len // 8)
            nt = hbseed.blastx_nt(7810 + k, nt_len, res[offs[5]:offs[5] + seg],
                                  res[offs[9]:offs[9] + seg])
            base = hbseed.blastx_query(nt)
        else:
            qlen = int(rng.choice(QLENS))
            base = hbseed.blastp_query(hbseed.random_protein(7810 + k, qlen))
        n_ctx = len(base.ctx_off)
        specs.append(hbseed.QuerySpec(base.seq, base.ctx_off, base.ctx_len,
                                      xdrop=rng.integers(5, 41, size=n_ctx).tolist(),
                                      cutoff=rng.integers(18, 61, size=n_ctx).tolist()))
    db = hbseed.Database(ctx, res, offs)
    b = hbseed.Batch(ctx, db, group_residues=512)
    try:
        ids = [b.add(s) for s in specs]
        b.run()
        for qid, spec in zip(ids, specs):
            _, exp = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=8)
            got = b.results(qid).view(np.int32).reshape(-1, 8)
            assert np.array_equal(got, exp.view(np.int32).reshape(-1, 8)), (qid, len(got), len(exp))
    finally:
        b.close()
        db.close()


@pytest.mark.parametrize("i", range(40))
def test_random_gapped_stage_matches_oracle(ctx, i):
    """The gapped stage on the HSPs of random scans, with random gap open
    (4-15), gap extension (1-3), X-drop (0-80) and per-context cutoffs."""
    from test_gpu_gapped import run_both
    spec, res, offs, _, _ = random_case(1000 + i)
    rng = np.random.default_rng(7900 + i)
    go, ge, x = int(rng.integers(4, 16)), int(rng.integers(1, 4)), int(rng.integers(0, 81))
    cut = rng.integers(25, 90, size=len(spec.ctx_off)).tolist()
    print(f"gapped case {i}: query {len(spec.seq)} residues, {len(offs) - 1} subjects, "
          f"open {go} extend {ge} X {x}", flush=True)
    run_both(ctx, spec, res, offs, cut, go, ge, x)
