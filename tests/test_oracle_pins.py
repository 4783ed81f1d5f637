"""Pins of the CPU oracle (CPU only, no GPU).

The reference ships no tests or golden vectors for this path and its CPU
word finder (NCBI BLAST+ 2.2.28 aa_ungapped.c) is not vendored, so each part
of the oracle is pinned here by something independent of its own code:
published tables, brute force over all words, closed forms of the X-drop
walks on designed score paths, hand-built two-hit cases, and a
diagonal-major restatement of the scan written differently from the
oracle's position-major loop.
"""
import os
import re

import numpy as np
import pytest

import hbseed
import oracle_lib

STDAA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
CODE = {c: i for i, c in enumerate(STDAA)}

# BLOSUM62 as published (Henikoff & Henikoff 1992), NCBI text order.
PUB_ORDER = "ARNDCQEGHILKMFPSTWYV"
PUB = """
 4 -1 -2 -2  0 -1 -1  0 -2 -1 -1 -1 -1 -2 -1  1  0 -3 -2  0
-1  5  0 -2 -3  1  0 -2  0 -3 -2  2 -1 -3 -2 -1 -1 -3 -2 -3
-2  0  6  1 -3  0  0  0  1 -3 -3  0 -2 -3 -2  1  0 -4 -2 -3
-2 -2  1  6 -3  0  2 -1 -1 -3 -4 -1 -3 -3 -1  0 -1 -4 -3 -3
 0 -3 -3 -3  9 -3 -4 -3 -3 -1 -1 -3 -1 -2 -3 -1 -1 -2 -2 -1
-1  1  0  0 -3  5  2 -2  0 -3 -2  1  0 -3 -1  0 -1 -2 -1 -2
-1  0  0  2 -4  2  5 -2  0 -3 -3  1 -2 -3 -1  0 -1 -3 -2 -2
 0 -2  0 -1 -3 -2 -2  6 -2 -4 -4 -2 -3 -3 -2  0 -2 -2 -3 -3
-2  0  1 -1 -3  0  0 -2  8 -3 -3 -1 -2 -1 -2 -1 -2 -2  2 -3
-1 -3 -3 -3 -1 -3 -3 -4 -3  4  2 -3  1  0 -3 -2 -1 -3 -1  3
-1 -2 -3 -4 -1 -2 -3 -4 -3  2  4 -2  2  0 -3 -2 -1 -2 -1  1
-1  2  0 -1 -3  1  1 -2 -1 -3 -2  5 -1 -3 -1  0 -1 -3 -2 -2
-1 -1 -2 -3 -1  0 -2 -3 -2  1  2 -1  5  0 -2 -1 -1 -1 -1  1
-2 -3 -3 -3 -2 -3 -3 -3 -1  0  0 -3  0  6 -4 -2 -2  1  3 -1
-1 -2 -2 -1 -3 -1 -1 -2 -2 -3 -3 -1 -2 -4  7 -1 -1 -4 -3 -2
 1 -1  1  0 -1  0  0  0 -1 -2 -2  0 -1 -2 -1  4  1 -3 -2 -2
 0 -1  0 -1 -1 -1 -1 -2 -2 -1 -1 -1 -1 -2 -1  1  5 -2 -2  0
-3 -3 -4 -4 -2 -2 -3 -2 -2 -3 -2 -3 -1  1 -4 -3 -2 11  2 -3
-2 -2 -2 -3 -2 -1 -2 -3  2 -1 -1 -2 -1  3 -3 -2 -2  2  7 -1
 0 -3 -3 -3 -1 -2 -2 -3 -3  3  1 -2  1 -1 -2 -2  0 -3 -1  4
"""
REF_MATRIX_FILE = "/root/reference/h_blast_v1.1/H-BLAST/h_blast.c"


def published_blosum62():
    vals = np.array(PUB.split(), dtype=np.int64).reshape(20, 20)
    return {(a, b): int(vals[i, j]) for i, a in enumerate(PUB_ORDER)
            for j, b in enumerate(PUB_ORDER)}


def test_oracle_matrix_standard_block_is_published_blosum62():
    m = oracle_lib.blosum62()
    pub = published_blosum62()
    for (a, b), v in pub.items():
        assert m[CODE[a], CODE[b]] == v, (a, b)
    assert np.array_equal(m, m.T)
    assert (m[0, :] == -128).all() and (m[:, 0] == -128).all()
    assert m[CODE["*"], CODE["*"]] == 1
    for a in PUB_ORDER:
        assert m[CODE["*"], CODE[a]] == -4


def test_product_and_oracle_matrices_agree():
    assert np.array_equal(hbseed.blosum62(), oracle_lib.blosum62())


@pytest.mark.skipif(not os.path.exists(REF_MATRIX_FILE), reason="reference tree not mounted")
def test_matrix_matches_reference_table():
    src = open(REF_MATRIX_FILE).read()
    body = src[src.index("Blosum62[28][32]"):]
    rows = re.findall(r"\{([-\d,\s]+)\}", body)[:28]
    ref = np.array([[int(x) for x in r.split(",")][:28] for r in rows])
    assert np.array_equal(oracle_lib.blosum62().astype(np.int64), ref)


# ---------------------------------------------------------------- lookup table
def brute_lut(seq, ctx_off, ctx_len, m, T, mask=None):
    """All 28^3 words scored at once against each query word (numpy); a word
    covering a masked residue is left out."""
    m = m.astype(np.int64)
    cells = {}
    for o, l in zip(ctx_off, ctx_len):
        for q in range(o, o + l - 2):
            if mask is not None and mask[q:q + 3].any():
                continue
            a, b, c = seq[q], seq[q + 1], seq[q + 2]
            sc = m[a][:, None, None] + m[b][None, :, None] + m[c][None, None, :]
            hit = sc >= T
            hit[a, b, c] = True
            for x, y, z in zip(*np.nonzero(hit)):
                cells.setdefault((int(x) << 10) | (int(y) << 5) | int(z), []).append(q)
    return cells


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lut_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    q = hbseed.random_protein(seed, 60)
    q[rng.integers(0, 57, 6)] = rng.choice([2, 21, 23, 24, 25, 26, 27], 6)  # ambiguity letters
    spec = hbseed.QuerySpec(q, [0, 25], [20, 35])
    spec.seq[20:25] = 0
    oq = oracle_lib.OracleQuery.from_spec(spec)
    start, ent = oq.lut_arrays()
    exp = brute_lut(spec.seq, spec.ctx_off, spec.ctx_len, oracle_lib.blosum62(), 11)
    got = {c: list(ent[start[c]:start[c + 1]]) for c in range(32768) if start[c + 1] > start[c]}
    assert got == exp


@pytest.mark.parametrize("seed", [4, 5])
def test_masked_lut_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    q = hbseed.random_protein(seed, 80)
    mask = np.zeros(80, np.uint8)
    for a in rng.integers(0, 75, 4):   # a few masked runs, e.g. SEG segments
        mask[a:a + int(rng.integers(1, 9))] = 1
    spec = hbseed.QuerySpec(q, [0, 35], [30, 45], lut_mask=mask)
    spec.seq[30:35] = 0
    oq = oracle_lib.OracleQuery.from_spec(spec)
    start, ent = oq.lut_arrays()
    exp = brute_lut(spec.seq, spec.ctx_off, spec.ctx_len, oracle_lib.blosum62(), 11, mask)
    got = {c: list(ent[start[c]:start[c + 1]]) for c in range(32768) if start[c + 1] > start[c]}
    assert got == exp
    # an all-masked query seeds nothing; masking nothing changes nothing
    assert oracle_lib.OracleQuery(q, [0], [80], [16], [41], lut_mask=np.ones(80)).lut.n_entries == 0
    a = oracle_lib.OracleQuery(q, [0], [80], [16], [41], lut_mask=np.zeros(80)).lut_arrays()
    b = oracle_lib.OracleQuery(q, [0], [80], [16], [41]).lut_arrays()
    assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_self_word_below_threshold_still_seeds():
    q = np.array([CODE[c] for c in "AXXXA"], np.uint8)   # XXX self-score -3 < 11
    oq = oracle_lib.OracleQuery.from_spec(hbseed.blastp_query(q))
    start, ent = oq.lut_arrays()
    xxx = (21 << 10) | (21 << 5) | 21
    assert list(ent[start[xxx]:start[xxx + 1]]) == [1]


def test_lookup_hit_counts_match_bruteforce():
    q = hbseed.random_protein(7, 80)
    res, offs = hbseed.make_db(8, 40, median=60, lo=1, hi=200)
    m = oracle_lib.blosum62().astype(np.int64)
    cnt, _ = oracle_lib.OracleQuery.from_spec(hbseed.blastp_query(q)).scan(res, offs)
    qw = np.stack([q[:-2], q[1:-1], q[2:]], 1).astype(np.int64)
    for i in range(len(offs) - 1):
        s = res[offs[i]:offs[i + 1]].astype(np.int64)
        if len(s) < 3:
            assert cnt[i, 0] == 0
            continue
        sw = np.stack([s[:-2], s[1:-1], s[2:]], 1)
        sc = (m[qw[:, None, 0], sw[None, :, 0]] + m[qw[:, None, 1], sw[None, :, 1]]
              + m[qw[:, None, 2], sw[None, :, 2]])
        same = (qw[:, None, :] == sw[None, :, :]).all(-1)
        assert cnt[i, 0] == int(((sc >= 11) | same).sum())


# ---------------------------------------------------------------- X-drop walks
def path_matrix(steps):
    """A matrix and two sequences whose aligned pairs score exactly `steps`."""
    m = np.full((28, 28), -1, np.int8)
    m[0, :] = m[:, 0] = -128
    vals = sorted(set(int(x) for x in steps))
    assert len(vals) <= 26
    letter = {v: i + 1 for i, v in enumerate(vals)}
    for v, L in letter.items():
        m[L, L] = v
    seq = np.array([letter[int(x)] for x in steps], np.uint8)
    return m, seq


def closed_form_right(steps, start_max, X):
    """Right walk by prefix sums: stop at the first i with S_i <= 0 or
    max(start, S_1..S_i) - S_i >= X; best = first index of the running max."""
    S = start_max + np.cumsum(steps)
    run = np.maximum.accumulate(np.maximum(S, start_max))
    stop = np.nonzero((S <= 0) | (run - S >= X))[0]
    last = stop[0] if len(stop) else len(steps)
    seen = S[:last + 1] if len(stop) else S
    best = -1
    if len(seen) and seen.max() > start_max:
        best = int(np.argmax(seen))
    return (int(max(start_max, seen.max() if len(seen) else start_max)), best + 1, int(last))


@pytest.mark.parametrize("seed", range(6))
def test_right_walk_closed_form(seed):
    rng = np.random.default_rng(seed)
    steps = rng.integers(-6, 7, 80)
    steps[:5] = [5, 4, 3, -2, 6]
    m, seq = path_matrix(steps)
    for X in (5, 10, 16):
        sc, ln, last = oracle_lib.extend_right(m, seq, seq, 0, 0, X, 0)
        esc, eln, elast = closed_form_right(steps, 0, X)
        assert (sc, ln, last) == (esc, eln, elast), (X, sc, ln, last, esc, eln, elast)


@pytest.mark.parametrize("seed", range(6))
def test_left_walk_closed_form(seed):
    rng = np.random.default_rng(100 + seed)
    steps = rng.integers(-6, 7, 60)
    m, seq = path_matrix(steps)
    n = len(steps) - 1
    for X in (5, 16):
        sc, ln = oracle_lib.extend_left(m, seq, seq, n, n, X, 0)
        # the left walk visits positions n, n-1, ..., 0 and has no score<=0 stop
        S = np.cumsum(steps[::-1])
        run = np.maximum.accumulate(np.maximum(S, 0))
        stop = np.nonzero(run - S >= X)[0]
        upto = stop[0] + 1 if len(stop) else len(S)
        seen = S[:upto]
        exp_len = int(np.argmax(seen)) + 1 if seen.max() > 0 else 0
        assert (sc, ln) == (max(0, int(seen.max())), exp_len)


def test_left_walk_at_offset_zero_follows_ncbi():
    """With n = min(s_off, q_off) = 0 the oracle walks position 0, as NCBI's
    loop (for i = n; i >= 0; i--) does; the reference kernel guards its walk
    with if (n > 0) (h_blast_kernel.cu:1368) and would return length 0."""
    m, seq = path_matrix([5])
    assert oracle_lib.extend_left(m, seq, seq, 0, 0, 16, 0) == (5, 1)
    m, seq = path_matrix([-3])
    assert oracle_lib.extend_left(m, seq, seq, 0, 0, 16, 0) == (0, 0)


def left_walk_starts(spec, res, offs, m, T=11, window=40):
    """min(s_off, q_off) at which every extension's left walk starts, from
    the two-hit rule restated per diagonal: the walk starts one residue before
    the trigger word's best prefix end (q + right_d - 1, s + right_d - 1)."""
    mi = m.astype(np.int64)
    seq = spec.seq
    words = []
    for o, l in zip(spec.ctx_off, spec.ctx_len):
        words.extend(range(o, o + l - 2))
    starts = []
    for oid in range(len(offs) - 1):
        s = res[offs[oid]:offs[oid + 1]]
        hl = []
        for b in range(len(s) - 2):
            for a in words:
                sc = sum(int(mi[seq[a + i], s[b + i]]) for i in range(3))
                if sc >= T or all(seq[a + i] == s[b + i] for i in range(3)):
                    hl.append((a, b))
        by_diag = {}
        for a, b in hl:
            by_diag.setdefault(a - b, []).append((a, b))
        for hits in by_diag.values():
            last, flag = None, False
            for a, b in hits:
                if flag:
                    if b >= last:
                        last, flag = b, False
                    continue
                if last is None or b - last >= window:
                    last = b
                    continue
                if b - last < 3:
                    continue
                sc, best, rd = 0, 0, 0
                for i in range(3):
                    sc += int(mi[seq[a + i], s[b + i]])
                    if sc > best:
                        best, rd = sc, i + 1
                starts.append(min(a + rd - 1, b + rd - 1))
                k = int(np.searchsorted(spec.ctx_off, a, side="right")) - 1
                r = oracle_lib.extend_two_hit(m, s, seq, last + 3, b, a, int(spec.xdrop[k]))
                if r["right"]:
                    last, flag = r["s_last"] - 2, True
                else:
                    last = b
    return starts


def test_left_walk_never_starts_below_offset_two():
    """Every extension is triggered by a hit at least 3 residues after an
    earlier hit of its diagonal on both sequences, so its left walk starts
    at offset >= 2 of both: the n = 0 case of the walk (where the reference
    kernel and NCBI differ) is never reached.  Brute force over repetitive
    queries and subjects that put words at offset 0 of both."""
    m = oracle_lib.blosum62()
    rng = np.random.default_rng(77)
    starts = []
    for t in range(6):
        unit = hbseed.random_protein(700 + t, 4)
        q = np.tile(unit, 8)[:30].astype(np.uint8)
        subj = [np.tile(unit, 6)[:20 + 3 * k].astype(np.uint8) for k in range(4)]
        subj.append(rng.choice(q, 25).astype(np.uint8))
        res = np.concatenate(subj)
        offs = np.zeros(len(subj) + 1, np.uint64)
        offs[1:] = np.cumsum([len(x) for x in subj])
        starts += left_walk_starts(hbseed.blastp_query(q), res, offs, m)
    # runs of X (BLOSUM62 X-X = -1: exact words seed with no improving
    # prefix, so the walk starts one residue before the trigger)
    X = CODE["X"]
    q = np.array([X] * 10 + list(hbseed.random_protein(790, 12)), np.uint8)
    s = np.array([X] * 12 + list(hbseed.random_protein(791, 10)), np.uint8)
    starts += left_walk_starts(hbseed.blastp_query(q), s, np.array([0, len(s)], np.uint64), m)
    assert len(starts) > 50
    assert min(starts) >= 2
    assert 2 in starts   # the bound is reached


def smallest_left_walk_case():
    """A two-hit case at the smallest offsets: Y B B B B B on both sequences
    with M[Y][Y] = 7, M[B][B] = -1 (exact words always seed), the rest -2.
    The trigger (3, 3) has no improving prefix (right_d = 0), so the left
    walk starts at offset 2 and walks 2, 1, 0: running -1, -2, +5 -> left
    length 3, score 5; the right walk from (3, 3) stops at once (score <= 0):
    HSP q_start = s_start = 0, length 3, score 5."""
    Y, B, A, C = CODE["Y"], CODE["B"], CODE["A"], CODE["C"]
    m = np.full((28, 28), -2, np.int8)
    m[0, :] = m[:, 0] = -128
    m[Y, Y] = 7
    m[B, B] = -1
    q = np.array([Y, B, B, B, B, B] + [A] * 30, np.uint8)
    s = np.array([Y, B, B, B, B, B] + [C] * 30, np.uint8)
    return m, q, s


def test_smallest_left_walk_crosses_offset_zero():
    m, q, s = smallest_left_walk_case()
    spec = hbseed.QuerySpec(q, [0], [len(q)], [16], [5])
    cnt, h = oracle_lib.OracleQuery.from_spec(spec, matrix=m).scan(
        s, np.array([0, len(s)], np.uint64))
    first = h[h["q_off"] == 3]
    assert len(first) == 1
    r = first[0]
    assert (r["q_start"], r["s_start"], r["s_off"], r["len"], r["score"]) == (0, 0, 3, 3, 5)


def test_right_walk_ties_keep_first_best():
    m, seq = path_matrix([3, -1, 1, 2, -16])  # running 3,2,3,5 -> peak 5 at index 3
    assert oracle_lib.extend_right(m, seq, seq, 0, 0, 16, 0) == (5, 4, 4)
    m, seq = path_matrix([3, -2, 2, -16])      # 3,1,3: equal peaks, first kept
    assert oracle_lib.extend_right(m, seq, seq, 0, 0, 16, 0) == (3, 1, 3)


def test_identical_segment_full_length():
    q = hbseed.random_protein(11, 70)
    m = oracle_lib.blosum62()
    self_score = int(sum(int(m[c, c]) for c in q))
    sc, ln, last = oracle_lib.extend_right(m, q, q, 0, 0, 16, 0)
    assert (sc, ln, last) == (self_score, 70, 70)
    sc, ln = oracle_lib.extend_left(m, q, q, 69, 69, 16, 0)
    assert (sc, ln) == (self_score, 70)


def test_sentinel_stops_extension():
    q = hbseed.random_protein(12, 30).copy()
    s = q.copy()
    q[15] = 0   # context sentinel inside the query
    m = oracle_lib.blosum62()
    sc, ln, last = oracle_lib.extend_right(m, s, q, 0, 0, 16, 0)
    assert ln == 15 and last == 15


# ---------------------------------------------------------------- two-hit rule
def two_word_case(k, extra=None):
    """Query and subject with 'WWW' words so that the only hits lie on one
    diagonal, at subject offsets 20 and 20 + k (query offsets 10, 10 + k)."""
    m = np.full((28, 28), -2, np.int8)
    m[0, :] = m[:, 0] = -128
    W, A, C = CODE["W"], CODE["A"], CODE["C"]
    m[W, W] = 6
    q = np.full(120, A, np.uint8)
    s = np.full(140, C, np.uint8)
    for off in (0, k):
        q[10 + off:13 + off] = W
        s[20 + off:23 + off] = W
    if extra is not None:
        extra(q, s)
    return m, q, s


@pytest.mark.parametrize("k,expect", [(1, 0), (2, 0), (3, 1), (20, 1), (39, 1), (40, 0), (60, 0)])
def test_two_hit_distance_rule(k, expect):
    m, q, s = two_word_case(k)
    spec = hbseed.QuerySpec(q, [0], [len(q)], [16], [1])
    oq = oracle_lib.OracleQuery.from_spec(spec, threshold=18, matrix=m)
    cnt, h = oq.scan(s, np.array([0, len(s)], np.uint64))
    hits_expected = 2 if k >= 3 else 2 + 0  # overlapping W runs give more words
    assert cnt[0, 1] == expect
    assert cnt[0, 0] >= hits_expected


def test_flagged_diagonal_skips_until_extension_end():
    # One diagonal (q - s = -10) carries W-triplet hits at s = 20, 25, 45, 90,
    # 95.  K runs (K-K = 5, KKK = 15 < T) make the extension from the pair
    # (20, 25) run right to s ~ 79 without creating words, so the hit at 45
    # lies inside the extended region and is skipped; 90 starts afresh and
    # 95 pairs with it.  Exactly two extensions.  The runs are K in the
    # query against R in the subject (K-R = 5) so that no exact word repeats
    # (an exact word seeds even below T).
    m = np.full((28, 28), -2, np.int8)
    m[0, :] = m[:, 0] = -128
    W, K, R, A, C = CODE["W"], CODE["K"], CODE["R"], CODE["A"], CODE["C"]
    m[W, W] = 6
    m[K, R] = m[R, K] = 5
    q = np.full(120, A, np.uint8)
    s = np.full(130, C, np.uint8)
    q[13:61] = K
    s[23:71] = R
    for qo in (10, 15, 35, 80, 85):
        q[qo:qo + 3] = W
        s[qo + 10:qo + 13] = W
    spec = hbseed.QuerySpec(q, [0], [len(q)], [16], [1])
    oq = oracle_lib.OracleQuery.from_spec(spec, threshold=18, matrix=m)
    cnt, h = oq.scan(s, np.array([0, len(s)], np.uint64))
    assert cnt[0, 0] == 25                     # every W triplet pair is a hit
    ext = h[h["q_off"] - h["s_off"] == -10]
    assert list(ext["s_off"]) == [25, 95]
    assert ext[0]["s_start"] == 20 and ext[0]["len"] >= 50
    # the other triplet pairs put two more pairs 5 apart on diagonals -80
    # (s = 90, 95) and 60 (s = 20, 25): 4 extensions in all
    assert cnt[0, 1] == 4
    other = h[h["q_off"] - h["s_off"] != -10]
    assert sorted(zip((other["q_off"] - other["s_off"]).tolist(), other["s_off"].tolist())) == \
        [(-80, 95), (60, 25)]
    ecnt, eh = diagonal_major_scan(spec, s, np.array([0, len(s)], np.uint64), T=18, m=m)
    assert ecnt[0, 1] == 4


# ---------------------------------------------------------------- scan restated
def diagonal_major_scan(spec, res, offs, T=11, window=40, m=None):
    """Same rules as the CPU word finder, organised per true diagonal
    (q - s) instead of per subject position, with the oracle's extension
    routine for the X-drop walks.  Returns (counters, sorted HSP tuples)."""
    m = oracle_lib.blosum62() if m is None else m
    mi = m.astype(np.int64)
    seq = spec.seq
    words = []
    for o, l in zip(spec.ctx_off, spec.ctx_len):
        words.extend(range(o, o + l - 2))
    words = np.array(words, np.int64)
    qw = np.stack([seq[words], seq[words + 1], seq[words + 2]], 1).astype(np.int64)
    counters, out = [], []
    for oid in range(len(offs) - 1):
        s = res[offs[oid]:offs[oid + 1]]
        if len(s) < 3:
            counters.append((0, 0, 0))
            continue
        sw = np.stack([s[:-2], s[1:-1], s[2:]], 1).astype(np.int64)
        sc = (mi[qw[:, None, 0], sw[None, :, 0]] + mi[qw[:, None, 1], sw[None, :, 1]]
              + mi[qw[:, None, 2], sw[None, :, 2]])
        hit = (sc >= T) | (qw[:, None, :] == sw[None, :, :]).all(-1)
        qi, si = np.nonzero(hit)
        qo, so = words[qi], si
        by_diag = {}
        for a, b in sorted(zip(qo.tolist(), so.tolist()), key=lambda t: (t[1], t[0])):
            by_diag.setdefault(a - b, []).append((a, b))
        n_ext = 0
        for d, hl in by_diag.items():
            last, flag = None, False
            for a, b in hl:
                if flag:
                    if b < last:
                        continue
                    last, flag = b, False
                    continue
                if last is None or b - last >= window:
                    last = b
                    continue
                if b - last < 3:
                    continue
                k = int(np.searchsorted(spec.ctx_off, a, side="right")) - 1
                r = oracle_lib.extend_two_hit(m, s, seq, last + 3, b, a, int(spec.xdrop[k]))
                n_ext += 1
                if r["score"] >= spec.cutoff[k]:
                    out.append((oid, k, r["q_start"], r["s_start"], a, b, r["len"], r["score"]))
                if r["right"]:
                    last, flag = r["s_last"] - 2, True
                else:
                    last = b
        n_hsp = sum(1 for o in out if o[0] == oid)
        counters.append((int(hit.sum()), n_ext, n_hsp))
    return np.array(counters, np.int64), sorted(out, key=lambda t: (t[0], t[5], t[4]))


@pytest.mark.parametrize("seed", [3, 4])
def test_oracle_scan_equals_diagonal_major_restatement(seed):
    q = hbseed.random_protein(seed, 120)
    res, offs = hbseed.make_db(seed + 10, 60, median=150, lo=1, hi=600, plant_query=q,
                               plant_frac=0.5)
    spec = hbseed.blastp_query(q)
    cnt, h = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs)
    ecnt, eh = diagonal_major_scan(spec, res, offs)
    assert np.array_equal(cnt.astype(np.int64), ecnt)
    assert [tuple(int(x) for x in r) for r in h] == eh
    assert len(eh) > 0


def test_oracle_multi_context_equals_restatement():
    # three contexts separated by sentinels (BLASTX-style concatenation)
    parts = [hbseed.random_protein(s, n) for s, n in ((21, 50), (22, 70), (23, 40))]
    seq = np.concatenate([parts[0], [0], parts[1], [0], parts[2]]).astype(np.uint8)
    spec = hbseed.QuerySpec(seq, [0, 51, 122], [50, 70, 40], [16, 12, 20], [30, 41, 35])
    res, offs = hbseed.make_db(24, 50, median=150, lo=1, hi=500, plant_query=seq,
                               plant_frac=0.6)
    cnt, h = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs)
    ecnt, eh = diagonal_major_scan(spec, res, offs)
    assert np.array_equal(cnt.astype(np.int64), ecnt)
    assert [tuple(int(x) for x in r) for r in h] == eh


def test_hsp_property_check_catches_planted_mistakes():
    """The property check used on results too large for the oracle passes
    the oracle's own records and flags one planted mistake of each kind: a
    score off by one, a start shifted along the subject (diagonal broken), a
    record past its subject's end, a record crossing a context boundary,
    a score below the cutoff and two records out of order."""
    q = hbseed.random_protein(41, 300)
    res, offs = hbseed.make_db(42, 2000, plant_query=q, plant_frac=0.2)
    spec = hbseed.blastp_query(q)
    oq = oracle_lib.OracleQuery.from_spec(spec)
    _, h = oq.scan(res, offs, threads=4)
    assert len(h) > 20
    assert oq.check_hsps(res, offs, h) == (0, len(h))
    slen = (offs[1:] - offs[:-1]).astype(np.int64)
    cases = []
    a = h.copy(); a["score"][3] += 1; cases.append((a, 3))
    a = h.copy(); a["s_start"][4] += 1; cases.append((a, 4))
    a = h.copy(); a["len"][5] = slen[a["oid"][5]] - a["s_start"][5] + 1; cases.append((a, 5))
    a = h.copy(); a["q_start"][6] = -1; cases.append((a, 6))
    a = h.copy(); a[[7, 8]] = a[[8, 7]]; cases.append((a, 8))
    for bad, idx in cases:
        n, first = oq.check_hsps(res, offs, bad)
        assert n >= 1 and first == idx, (idx, n, first)
    hi = oracle_lib.OracleQuery(spec.seq, spec.ctx_off, spec.ctx_len, spec.xdrop,
                                [int(h["score"].min()) + 1])
    n, _ = hi.check_hsps(res, offs, h)
    assert n == int((h["score"] < int(h["score"].min()) + 1).sum()) > 0


def test_hsp_invariants():
    q = hbseed.random_protein(31, 300)
    res, offs = hbseed.make_db(32, 3000, plant_query=q, plant_frac=0.1)
    cnt, h = oracle_lib.OracleQuery.from_spec(hbseed.blastp_query(q)).scan(res, offs, threads=4)
    assert (cnt[:, 2] <= cnt[:, 1]).all() and (cnt[:, 1] <= cnt[:, 0]).all()
    assert ((h["q_off"] - h["s_off"]) == (h["q_start"] - h["s_start"])).all()
    assert (h["score"] >= 41).all()
    m = oracle_lib.blosum62().astype(np.int64)
    for r in h[:200]:
        b = int(offs[int(r["oid"])]) + int(r["s_start"])
        s = res[b:b + int(r["len"])]
        qq = q[int(r["q_start"]):int(r["q_start"]) + int(r["len"])]
        assert int(m[qq, s].sum()) == r["score"]
        assert r["s_start"] <= r["s_off"] + 3 and r["s_off"] <= r["s_start"] + r["len"]


# ---------------------------------------------------------------- six frames
CODONS = {
    "F": "TTT TTC", "L": "TTA TTG CTT CTC CTA CTG", "I": "ATT ATC ATA", "M": "ATG",
    "V": "GTT GTC GTA GTG", "S": "TCT TCC TCA TCG AGT AGC", "P": "CCT CCC CCA CCG",
    "T": "ACT ACC ACA ACG", "A": "GCT GCC GCA GCG", "Y": "TAT TAC", "*": "TAA TAG TGA",
    "H": "CAT CAC", "Q": "CAA CAG", "N": "AAT AAC", "K": "AAA AAG", "D": "GAT GAC",
    "E": "GAA GAG", "C": "TGT TGC", "W": "TGG", "R": "CGT CGC CGA CGG AGA AGG",
    "G": "GGT GGC GGA GGG",
}
TABLE = {c: aa for aa, cs in CODONS.items() for c in cs.split()}


def test_codon_table_complete():
    assert len(TABLE) == 64


def test_six_frame_translation():
    rng = np.random.default_rng(5)
    nt = "".join(rng.choice(list("ACGT"), 101))
    spec = hbseed.blastx_query(nt)
    comp = {"A": "T", "C": "G", "G": "C", "T": "A"}
    rc = "".join(comp[c] for c in reversed(nt))
    strands = [nt, nt, nt, rc, rc, rc]
    for f in range(6):
        src, start = strands[f], f % 3
        aa = "".join(TABLE[src[i:i + 3]] for i in range(start, len(src) - 2, 3))
        o, l = spec.ctx_off[f], spec.ctx_len[f]
        assert hbseed.decode_protein(spec.seq[o:o + l]) == aa
        if f:
            assert spec.seq[o - 1] == 0
    assert len(spec.seq) == sum(spec.ctx_len) + 5


def test_reverse_translation_round_trip():
    p = hbseed.random_protein(6, 100)
    nt = hbseed.blastx_nt(7, 600, p, p[::-1].copy())
    spec = hbseed.blastx_query(nt)
    text = hbseed.decode_protein(p)
    frames = [hbseed.decode_protein(spec.seq[o:o + l]) for o, l in zip(spec.ctx_off, spec.ctx_len)]
    assert any(text in f for f in frames[:3])
    assert any(hbseed.decode_protein(p[::-1]) in f for f in frames[3:])
