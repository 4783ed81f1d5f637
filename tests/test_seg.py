"""SEG low-complexity masking (hb_seg_mask): the oracle's entropy and
probability against their closed forms, the product's mask against the
oracle's on crafted and random queries, and the invariants the algorithm
fixes (no GPU).

NCBI's implementation (blast_seg.c) is not in the reference, so equality
with NCBI's own masks is not pinned here: both sides restate the published
algorithm (Wootton and Federhen 1993, 1996) with NCBI's -seg 12 2.2 2.5.
"""
import math

import numpy as np
import pytest

import oracle_lib as O

hb = pytest.importorskip("hbseed")

STD20 = [1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 22]
BOGUS = [2, 21, 23, 24, 25, 26, 27]          # B X Z U * O J


def counts_of(letters):
    c = np.zeros(28, np.int32)
    for x in letters:
        c[x] += 1
    return c


# --- closed forms ------------------------------------------------------------

def test_entropy_by_hand():
    assert O.seg_entropy(counts_of([1] * 12)) == 0.0
    assert abs(O.seg_entropy(counts_of([1] * 6 + [3] * 6)) - 1.0) < 1e-15
    assert abs(O.seg_entropy(counts_of(STD20[:12])) - math.log2(12)) < 1e-14
    assert abs(O.seg_entropy(counts_of([1] * 8 + [3] * 4)) -
               (-(2 / 3) * math.log2(2 / 3) - (1 / 3) * math.log2(1 / 3))) < 1e-15


def test_entropy_is_shannon_entropy():
    rng = np.random.default_rng(1)
    for _ in range(2000):
        c = np.zeros(28, np.int32)
        k = int(rng.integers(1, 21))
        letters = rng.choice(STD20, k, replace=False)
        c[letters] = rng.integers(1, 40, k)
        p = c[c > 0] / c.sum()
        assert abs(O.seg_entropy(c) - float(-(p * np.log2(p)).sum())) < 1e-12


def lnprob_closed_form(c, n):
    # ln( 20! / prod_k m_k! * n! / prod_i c_i! / 20^n ), m_k = letters with count k
    counts = [int(x) for x in c[STD20]]
    mult = {}
    for x in counts:
        mult[x] = mult.get(x, 0) + 1
    ass = math.lgamma(21) - sum(math.lgamma(m + 1) for m in mult.values())
    perm = math.lgamma(n + 1) - sum(math.lgamma(x + 1) for x in counts)
    return ass + perm - n * math.log(20)


def test_lnprob_by_hand_and_closed_form():
    # n copies of one letter: 20 of the 20^n sequences -> ln P = -(n - 1) ln 20
    for n in (1, 2, 12, 40):
        assert abs(O.seg_lnprob(counts_of([5] * n), n) + (n - 1) * math.log(20)) < 1e-10
    # two letters once each: 20 * 19 ordered pairs of distinct letters
    assert abs(O.seg_lnprob(counts_of([1, 3]), 2) - math.log(20 * 19 / 400)) < 1e-12
    rng = np.random.default_rng(2)
    for _ in range(2000):
        c = np.zeros(28, np.int32)
        k = int(rng.integers(1, 21))
        c[rng.choice(STD20, k, replace=False)] = rng.integers(1, 12, k)
        n = int(c.sum())
        assert abs(O.seg_lnprob(c, n) - lnprob_closed_form(c, n)) < 1e-9


def test_lnprob_sums_to_one_over_compositions():
    # over all compositions of n letters from 20, the probabilities sum to 1
    # (n = 4: the partitions of 4 give every composition class)
    total = 0.0
    for parts in ([4], [3, 1], [2, 2], [2, 1, 1], [1, 1, 1, 1]):
        c = counts_of(sum(([STD20[i]] * p for i, p in enumerate(parts)), []))
        total += math.exp(O.seg_lnprob(c, 4))
    assert abs(total - 1.0) < 1e-12


# --- product against the oracle ---------------------------------------------

BG = np.array([78, 19, 54, 63, 39, 74, 22, 51, 57, 90, 22, 45, 52, 43, 51, 71, 59, 64, 13, 32],
              float)


def background(rng, n):
    return rng.choice(STD20, n, p=BG / BG.sum()).astype(np.uint8)


def low_complexity(rng, n):
    kind = int(rng.integers(0, 4))
    if kind == 0:
        return np.full(n, rng.choice(STD20), np.uint8)
    if kind == 1:
        a, b = rng.choice(STD20, 2, replace=False)
        return np.array([a, b] * (n // 2 + 1), np.uint8)[:n]
    if kind == 2:
        letters = rng.choice(STD20, 3, replace=False)
        return rng.choice(letters, n).astype(np.uint8)
    x = background(rng, n)
    x[rng.random(n) < 0.5] = rng.choice(STD20)
    return x


def random_query(rng):
    parts = []
    for _ in range(int(rng.integers(1, 6))):
        parts.append(background(rng, int(rng.integers(0, 200))))
        if rng.random() < 0.7:
            parts.append(low_complexity(rng, int(rng.integers(3, 80))))
        if rng.random() < 0.2:
            parts.append(rng.choice(BOGUS, int(rng.integers(1, 6))).astype(np.uint8))
    return np.concatenate(parts) if parts else np.zeros(0, np.uint8)


def test_product_matches_oracle_random_queries():
    rng = np.random.default_rng(3)
    masked = 0
    for it in range(600):
        q = random_query(rng)
        spec = hb.blastp_query(q)
        a = hb.seg_mask(spec)
        b = O.seg_mask(q, [0], [q.size])
        assert np.array_equal(a, b), it
        masked += int(a.sum())
    assert masked > 1000


@pytest.mark.parametrize("window,locut,hicut,maxtrim", [
    (12, 2.2, 2.5, 100), (8, 1.5, 1.8, 100), (20, 3.0, 3.3, 100), (12, 2.2, 2.5, 5),
    (12, 0.0, 0.0, 100), (45, 3.4, 3.75, 100), (1, 0.0, 0.0, 100)])
def test_product_matches_oracle_other_parameters(window, locut, hicut, maxtrim):
    rng = np.random.default_rng(window * 100 + maxtrim)
    for it in range(150):
        q = random_query(rng)
        spec = hb.blastp_query(q)
        a = hb.seg_mask(spec, window=window, locut=locut, hicut=hicut, maxtrim=maxtrim)
        b = O.seg_mask(q, [0], [q.size], window, locut, hicut, 2, maxtrim)
        assert np.array_equal(a, b), (it, window, locut, hicut, maxtrim)


def test_blastx_contexts_are_masked_apart():
    # a CAG repeat (poly-Q in frame +1) between random codons: the frames
    # are masked one by one and no mask reaches a sentinel
    rng = np.random.default_rng(4)
    nt = "".join(rng.choice(list("ACGT"), 600)) + "CAG" * 40 + "".join(rng.choice(list("ACGT"), 600))
    spec = hb.blastx_query(nt)
    a = hb.seg_mask(spec)
    b = O.seg_mask(spec.seq, spec.ctx_off, spec.ctx_len)
    assert np.array_equal(a, b)
    assert a.sum() >= 3 * 30
    inside = np.zeros(spec.seq.size, bool)
    for o, n in zip(spec.ctx_off, spec.ctx_len):
        inside[o:o + n] = True
    assert not a[~inside].any()


# --- invariants ----------------------------------------------------------------

def high_complexity(rng, n):
    # blocks that are permutations of the 20 letters: a 12-window holds each
    # letter at most twice, so its entropy stays far above 2.5 bits
    return np.concatenate([rng.permutation(STD20) for _ in range(n // 20 + 1)])[:n].astype(
        np.uint8)


def test_homopolymer_run_is_masked_and_mask_stays_near_it():
    rng = np.random.default_rng(5)
    for it in range(100):
        L = int(rng.integers(12, 60))
        left, right = high_complexity(rng, 150), high_complexity(rng, 150)
        assert not hb.seg_mask(hb.blastp_query(np.concatenate([left, right]))).any()
        q = np.concatenate([left, np.full(L, rng.choice(STD20), np.uint8), right])
        m = hb.seg_mask(hb.blastp_query(q))
        assert m[150:150 + L].all(), it
        # the segment may take a few flanking residues, never far ones
        assert not m[:150 - 12].any() and not m[150 + L + 12:].any(), it


def test_no_mask_without_low_complexity():
    # cycling through all 20 letters: every window holds 12 distinct
    # letters (log2 12 = 3.58 bits > 2.5)
    q = np.array(STD20 * 10, np.uint8)
    assert not hb.seg_mask(hb.blastp_query(q)).any()
    # shorter than the window
    assert not hb.seg_mask(hb.blastp_query(np.full(11, 1, np.uint8))).any()
    assert hb.seg_mask(hb.blastp_query(np.full(12, 1, np.uint8))).all()
    # runs of non-standard letters get no entropy: X runs are not SEG's
    q = np.concatenate([background(np.random.default_rng(6), 100), np.full(40, 21, np.uint8),
                        background(np.random.default_rng(7), 100)])
    assert not hb.seg_mask(hb.blastp_query(q))[100:140].any()


def test_seg_query_hard_and_soft():
    q = np.concatenate([background(np.random.default_rng(8), 100), np.full(30, 16, np.uint8),
                        background(np.random.default_rng(9), 100)])
    spec = hb.blastp_query(q)
    m = hb.seg_mask(spec)
    hard = hb.seg_query(spec)
    assert np.array_equal(hard.lut_mask, m)
    assert (hard.seq[m != 0] == 21).all() and np.array_equal(hard.seq[m == 0], q[m == 0])
    soft = hb.seg_query(spec, hard=False)
    assert np.array_equal(soft.seq, q) and np.array_equal(soft.lut_mask, m)


def test_seg_rejects_bad_input():
    spec = hb.blastp_query(np.array([1, 2, 3], np.uint8))
    spec.seq[1] = 30
    with pytest.raises(hb.HBError) as e:
        hb.seg_mask(spec)
    assert e.value.code == -5
    with pytest.raises(hb.HBError):
        hb.seg_mask(hb.blastp_query(np.ones(5, np.uint8)), window=0)
    with pytest.raises(hb.HBError):
        hb.seg_mask(hb.blastp_query(np.ones(5, np.uint8)), locut=3.0, hicut=2.0)
