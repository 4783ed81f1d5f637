"""Pins of the oracle's preliminary gapped extension (CPU only, no GPU).

The reference vendors the gapped-extension source it calls
(ncbi_blast_files/blast_gapalign.c) but no tests or vectors for it, so the
oracle's restatement is pinned by things independent of its code:
  * at a drop-off too large to prune, the X-drop DP must equal a textbook
    affine-gap (Gotoh) DP anchored at the origin, written here as full
    matrices -- this fixes the recurrence, the gap costs and the convention
    for reporting the best cell;
  * for every drop-off, the banded DP must equal a full-matrix statement of
    the same pruning rule (a cell is dead when its score falls more than X
    below the best score seen so far in row-major order; dead cells feed
    nothing), plus the one property of the band the full matrix cannot
    express: row 0 is closed without a guard column, so the diagonal out of
    its last open column is not taken;
  * closed forms on designed sequences: a mismatch run that costs exactly X
    survives and one that costs X + 1 ends the extension (strict '>'), the
    drop-off is raised to gap_open + gap_extend, empty sides score 0;
  * the start point (best 11-residue window, first maximum, the window's
    middle; the HSP's middle when it is 11 residues or shorter; its start
    when no window scores above 0) on designed HSPs;
  * the two-sided extension equals left and right Gotoh DPs computed
    separately, and results do not change when a subject is padded far away
    (covers the long-subject window of AdjustSubjectRange);
  * containment on designed subjects, and invariants on random databases.
"""
import numpy as np
import pytest

import hbseed
import oracle_lib as O

STDAA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
CODE = {c: i for i, c in enumerate(STDAA)}
M62 = O.blosum62()
NEG = -(10 ** 9)


def enc(s):
    return np.array([CODE[c] for c in s], np.uint8)


def gotoh(a, b, go=11, ge=1):
    """Full affine DP anchored at the origin: best over cells (i >= 1) of the
    best alignment of a[:i] with b[:j]; first maximum in row-major order.
    Returns (score, i, j); (0, 0, 0) when no cell scores above 0."""
    M, N = len(a), len(b)
    H = np.full((M + 1, N + 1), NEG, np.int64)
    E = np.full((M + 1, N + 1), NEG, np.int64)
    F = np.full((M + 1, N + 1), NEG, np.int64)
    H[0, 0] = 0
    for j in range(1, N + 1):
        H[0, j] = E[0, j] = -(go + j * ge)
    for i in range(1, M + 1):
        H[i, 0] = F[i, 0] = -(go + i * ge)
    for i in range(1, M + 1):
        for j in range(1, N + 1):
            E[i, j] = max(H[i, j - 1] - go - ge, E[i, j - 1] - ge)
            F[i, j] = max(H[i - 1, j] - go - ge, F[i - 1, j] - ge)
            H[i, j] = max(H[i - 1, j - 1] + int(M62[a[i - 1], b[j - 1]]), E[i, j], F[i, j])
    best, bi, bj = 0, 0, 0
    for i in range(1, M + 1):
        for j in range(N + 1):
            if H[i, j] > best:
                best, bi, bj = int(H[i, j]), i, j
    return best, bi, bj


def pruned_full(a, b, X, go=11, ge=1):
    """The X-drop rule on the full matrix, cells visited in row-major order.
    A cell's inputs come from live cells only; it is live when its score is
    no more than X below the best score seen before it.  Row 0 holds the
    leading gaps that stay within X; the diagonal out of its last live
    column into row 1 is not taken (row 0 has no guard column)."""
    goe = go + ge
    X = max(X, goe)
    M, N = len(a), len(b)
    H = np.full((M + 1, N + 1), NEG, np.int64)
    V = np.full((M + 1, N + 1), NEG, np.int64)  # best ending in a gap in b, into next row
    H[0, 0], V[0, 0] = 0, -goe
    end0 = 1
    while end0 <= N and -(go + end0 * ge) >= -X:
        H[0, end0] = -(go + end0 * ge)
        V[0, end0] = H[0, end0] - goe
        end0 += 1
    top, bi, bj = 0, 0, 0
    for i in range(1, M + 1):
        R = NEG  # best ending in a gap in a, along the row
        any_live = False
        for j in range(N + 1):
            diag = NEG
            if j >= 1 and not (i == 1 and j == end0):
                diag = H[i - 1, j - 1] + int(M62[a[i - 1], b[j - 1]])
            s = max(diag, V[i - 1, j], R)
            if top - s > X or s <= NEG // 2:
                R = NEG
                continue
            any_live = True
            H[i, j] = s
            if s > top:
                top, bi, bj = int(s), i, j
            V[i, j] = max(s - goe, V[i - 1, j] - ge)
            R = max(s - goe, R - ge)
        if not any_live:
            break
    return top, bi, bj


def mutate(rng, src, n_out):
    """A homolog-like relative of src: substitutions and short indels."""
    out = []
    for c in src:
        r = rng.random()
        if r < 0.08:
            continue
        if r < 0.16:
            out.append(int(rng.integers(1, 23)))
        out.append(int(c) if rng.random() < 0.7 else int(rng.integers(1, 23)))
    out = np.array(out[:n_out] if out else [1], np.uint8)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_semigapped_equals_gotoh_without_pruning(seed):
    rng = np.random.default_rng(seed)
    for _ in range(30):
        a = rng.integers(1, 23, size=int(rng.integers(0, 24))).astype(np.uint8)
        b = mutate(rng, a, 26) if len(a) and rng.random() < 0.7 else \
            rng.integers(1, 23, size=int(rng.integers(0, 24))).astype(np.uint8)
        M, N = len(a), len(b)
        exp = gotoh(a, b)
        # forward: A[1..M], B[1..N] (index 0 unused)
        A = np.concatenate([[0], a]).astype(np.uint8)
        B = np.concatenate([[0], b]).astype(np.uint8)
        got = O.semigapped(M62, A, B, M, N, 0, x_dropoff=10 ** 6)
        assert got == (exp if M and N else (0, 0, 0)), (seed, a, b)
        # reverse: A[M-1..0], B[N-1..0]
        got_r = O.semigapped(M62, a[::-1].copy(), b[::-1].copy(), M, N, 1, x_dropoff=10 ** 6)
        assert got_r == got


@pytest.mark.parametrize("seed", range(8))
def test_banded_dp_equals_full_matrix_pruning(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(40):
        a = rng.integers(1, 23, size=int(rng.integers(1, 40))).astype(np.uint8)
        b = mutate(rng, a, 45) if rng.random() < 0.8 else \
            rng.integers(1, 23, size=int(rng.integers(1, 40))).astype(np.uint8)
        X = int(rng.choice([0, 5, 12, 13, 16, 20, 25, 38, 60]))
        go, ge = [(11, 1), (10, 1), (9, 2), (5, 2)][int(rng.integers(0, 4))]
        A = np.concatenate([[0], a]).astype(np.uint8)
        B = np.concatenate([[0], b]).astype(np.uint8)
        got = O.semigapped(M62, A, B, len(a), len(b), 0, go, ge, X)
        assert got == pruned_full(a, b, X, go, ge), (seed, X, go, ge)


def test_row_zero_diagonal_rule_is_observable():
    # A case where the diagonal out of row 0's last open column would matter:
    # b opens with 27 residues the pruning keeps only as leading gaps, then a
    # residue matching a's first.
    a = enc("W" + "AAAA")
    b = enc("G" * 27 + "W" + "AAAA")
    A = np.concatenate([[0], a]).astype(np.uint8)
    B = np.concatenate([[0], b]).astype(np.uint8)
    got = O.semigapped(M62, A, B, len(a), len(b), 0, 11, 1, 38)
    assert got == pruned_full(a, b, 38)


def flanks(mid_a, mid_b):
    left, right = "WCWCW", "CWCWCWCWCW"   # 51 and 100, no useful shifts
    return enc(left + mid_a + right), enc(left + mid_b + right)


@pytest.mark.parametrize("reverse", [0, 1])
def test_xdrop_boundary_is_strict(reverse):
    a, b = flanks("LLLLL", "DDDDD")       # five L/D pairs: -20
    if reverse:
        # the reverse DP reads A[M-1], ..., A[0]: store the sequences reversed
        A, B = a[::-1].copy(), b[::-1].copy()
    else:
        A = np.concatenate([[0], a]).astype(np.uint8)
        B = np.concatenate([[0], b]).astype(np.uint8)
    n = len(a)
    assert O.semigapped(M62, A, B, n, n, reverse, 11, 1, 20) == (131, n, n)
    assert O.semigapped(M62, A, B, n, n, reverse, 11, 1, 19) == (51, 5, 5)


def test_xdrop_raised_to_gap_open_extend():
    a, b = flanks("LLL", "DDD")           # -12 = gap_open + gap_extend
    A = np.concatenate([[0], a]).astype(np.uint8)
    B = np.concatenate([[0], b]).astype(np.uint8)
    n = len(a)
    assert O.semigapped(M62, A, B, n, n, 0, 11, 1, 3) == (139, n, n)
    assert O.semigapped(M62, A, B, n, n, 0, 11, 1, 12) == (139, n, n)


def test_empty_sides_score_zero():
    A = enc("-WWW")
    assert O.semigapped(M62, A, A, 0, 3, 0) == (0, 0, 0)
    assert O.semigapped(M62, A, A, 3, 0, 0) == (0, 0, 0)


def test_gap_cheaper_than_mismatches():
    # b carries 3 extra residues in the middle: one gap of 3 costs 14, the
    # shifted ungapped path costs far more, so the DP bridges with the gap.
    a = enc("WCWCWCWCWC" + "WCWCWCWCWC")
    b = enc("WCWCWCWCWC" + "GGG" + "WCWCWCWCWC")
    A = np.concatenate([[0], a]).astype(np.uint8)
    B = np.concatenate([[0], b]).astype(np.uint8)
    assert O.semigapped(M62, A, B, len(a), len(b), 0) == (200 - 14, 20, 23)


# ---------------------------------------------------------------- start point
def test_start_short_hsp_is_middle():
    q = enc("AAAAAAAAAAAAAAAA")
    assert O.gapped_start(M62, q, q, 2, 11, 2, 11) == 2 + 5
    assert O.gapped_start(M62, q, q, 3, 6, 3, 6) == 3 + 3


def test_start_is_middle_of_best_window():
    # 30-residue HSP: mismatching pairs except an 11-residue identical run of
    # W at offsets 12..22 -> best window ends at 22, start 22 - 5 = 17.
    q = enc("L" * 12 + "W" * 11 + "L" * 7)
    s = enc("D" * 12 + "W" * 11 + "D" * 7)
    assert O.gapped_start(M62, q, s, 0, 30, 0, 30) == 17


def test_start_first_maximum_wins():
    q = enc("W" * 11 + "L" * 5 + "W" * 11)
    s = enc("W" * 11 + "D" * 5 + "W" * 11)
    assert O.gapped_start(M62, q, s, 0, 27, 0, 27) == 10 - 5


def test_start_without_positive_window_is_hsp_start():
    q = enc("L" * 20)
    s = enc("D" * 20)
    assert O.gapped_start(M62, q, s, 0, 20, 0, 20) == 0


# ---------------------------------------------------------- two-sided extension
@pytest.mark.parametrize("seed", range(4))
def test_prot_gapped_equals_two_gotoh_halves(seed):
    rng = np.random.default_rng(200 + seed)
    for _ in range(15):
        q = rng.integers(1, 23, size=int(rng.integers(2, 30))).astype(np.uint8)
        s = mutate(rng, q, 34)
        if len(s) < 2:
            continue
        qo = int(rng.integers(0, len(q)))
        so = int(rng.integers(0, len(s)))
        got = O.prot_gapped(M62, q, s, qo, so, x_dropoff=10 ** 6)
        left = gotoh(q[:qo + 1][::-1], s[:so + 1][::-1])
        if qo + 1 < len(q) and so + 1 < len(s):
            right = gotoh(q[qo + 1:], s[so + 1:])
            qe, se = qo + 1 + right[1], so + 1 + right[2]
        else:
            right, qe, se = (0, 0, 0), qo, so
        exp = (left[0] + right[0], qo + 1 - left[1], qe, so + 1 - left[2], se)
        assert got == exp, (seed, qo, so)


def test_long_subject_window_keeps_results():
    rng = np.random.default_rng(7)
    q = rng.integers(1, 23, size=300).astype(np.uint8)
    core = mutate(rng, q[100:200], 100)
    short = np.concatenate([rng.integers(1, 23, size=50), core,
                            rng.integers(1, 23, size=50)]).astype(np.uint8)
    pad = 95000
    far = np.concatenate([rng.integers(1, 23, size=pad), short,
                          rng.integers(1, 23, size=pad)]).astype(np.uint8)
    qo, so = 150, 50 + 50
    a = O.prot_gapped(M62, q, short, qo, so)
    b = O.prot_gapped(M62, q, far, qo, so + pad)
    assert a[0] == b[0] and a[1:3] == b[1:3]
    assert (a[3] + pad, a[4] + pad) == b[3:5]
    # the subject window is exercised (subject >= 90000 residues)
    assert len(far) >= 90000


# ----------------------------------------------------------------- containment
def _gapped_of(q_text, subjects, hsps, cutoff=20):
    q = enc(q_text)
    res = np.concatenate(subjects).astype(np.uint8)
    offs = np.concatenate([[0], np.cumsum([len(x) for x in subjects])]).astype(np.uint64)
    oq = O.OracleQuery(q, [0], [len(q)], [16], [41])
    recs = np.zeros(len(hsps), O.HSP_DTYPE)
    for i, h in enumerate(hsps):
        recs[i] = h
    return O.gapped_db(oq, res, offs, recs, [cutoff])


def test_duplicate_hsp_is_contained():
    q = "WCWCWCWCWCWCWCWCWCWC"
    h = (0, 0, 0, 0, 5, 5, 20, 200)
    cnt, g = _gapped_of(q, [enc(q)], [h, h])
    assert cnt.tolist() == [[1, 1]]
    assert (g["score"][0], g["q_start"][0], g["q_end"][0]) == (200, 0, 20)


def test_separate_regions_both_extended():
    q = "WCWCWCWCWC"
    s = enc(q + "G" * 40 + q)
    hs = [(0, 0, 0, 0, 5, 5, 10, 100), (0, 0, 0, 50, 5, 55, 10, 100)]
    cnt, g = _gapped_of(q, [s], hs)
    assert cnt.tolist() == [[2, 2]]
    # ordered by score, then subject start
    assert list(g["s_start"]) == [0, 50]


def test_higher_scoring_inner_hsp_not_contained():
    q = "WCWCWCWCWC"
    hs = [(0, 0, 0, 0, 5, 5, 10, 100), (0, 0, 2, 2, 5, 5, 4, 101)]
    cnt, g = _gapped_of(q, [enc(q)], hs)
    # the inner one (101) goes first and is extended to the full run (100);
    # the outer one (100 <= 100) is then contained
    assert cnt.tolist() == [[1, 1]]
    # an HSP scoring above the saved gapped alignment that holds it is not
    # contained: the inner one (120) goes first and is saved with 100; the
    # outer one (110 > 100) is extended too
    hs = [(0, 0, 0, 0, 5, 5, 10, 110), (0, 0, 2, 2, 5, 5, 4, 120)]
    cnt, _ = _gapped_of(q, [enc(q)], hs)
    assert cnt.tolist() == [[2, 2]]


def test_below_cutoff_not_saved_and_not_containing():
    q = "WCWCWCWCWC"
    h = (0, 0, 0, 0, 5, 5, 10, 100)
    cnt, g = _gapped_of(q, [enc(q)], [h, h], cutoff=101)
    assert cnt.tolist() == [[0, 2]] and len(g) == 0


# ------------------------------------------------------------------ invariants
def test_gapped_invariants_on_random_database():
    qseq = hbseed.random_protein(31, 400)
    res, offs = hbseed.make_db(32, 400, plant_query=qseq, plant_frac=0.3)
    spec = hbseed.blastp_query(qseq)
    oq = O.OracleQuery.from_spec(spec)
    _, hsps = oq.scan(res, offs)
    cnt, g = O.gapped_db(oq, res, offs, hsps, [60], threads=4)
    assert len(g) == cnt[:, 0].sum() and len(g) > 10
    per = np.bincount(hsps["oid"], minlength=len(offs) - 1)
    assert np.all(cnt[:, 0] <= cnt[:, 1]) and np.all(cnt[:, 1] <= per)
    assert np.all(cnt[per > 0, 1] >= 1)     # the best HSP is always extended
    assert np.all(g["score"] >= 60)
    assert np.all(g["q_start"] <= g["q_gstart"]) and np.all(g["q_gstart"] < g["q_end"])
    assert np.all(g["s_start"] <= g["s_gstart"]) and np.all(g["s_gstart"] < g["s_end"])
    # every start point lies on the diagonal of one of its subject's HSPs
    for r in g:
        h = hsps[hsps["oid"] == r["oid"]]
        assert np.any(h["q_start"] - h["s_start"] == r["q_gstart"] - r["s_gstart"])
    slen = np.diff(offs.astype(np.int64))
    assert np.all(g["s_end"] <= slen[g["oid"]]) and np.all(g["q_end"] <= 400)
    # each saved HSP's score equals an unpruned-or-pruned DP over its own box:
    # re-extending from its start point reproduces it exactly
    for r in g[:20]:
        s = res[offs[r["oid"]]:offs[r["oid"] + 1]]
        got = O.prot_gapped(M62, qseq, s, int(r["q_gstart"]), int(r["s_gstart"]))
        assert got == (r["score"], r["q_start"], r["q_end"], r["s_start"], r["s_end"])
    # threads do not change the result
    cnt1, g1 = O.gapped_db(oq, res, offs, hsps, [60], threads=1)
    assert np.array_equal(cnt, cnt1) and np.array_equal(g.view(np.int32), g1.view(np.int32))
