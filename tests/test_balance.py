"""CPU/GPU split of a volume (load balancing): the options file, the split of
a volume's .hDb at a workload ratio and the ratio's update from wall times
(ncbi_blast_files/functions.cpp:142-203, 449-571, 51-140), product against
the oracle and both against hand-worked values, an independent brute-force
statement of the split and the invariants the formulas fix (no GPU).
"""
import numpy as np
import pytest

import oracle_lib as O

hb = pytest.importorskip("hbseed")

OPTS_TEXT = ("#GPU db used ratio\n0.7\n#fixed\n0\n#capability\n10.5\n#max hits\n30\n"
             "#ecc\n1\n#balancer\n0\n#load factor\n6\n#chunk\n150\n#base value\n80\n")


# --- options file (functions.cpp:142-203) -----------------------------------

def test_options_defaults_by_hand():
    want = dict(gpu_ratio=-1.0, fixed_ratio=1, capability_rate=12.0, max_hits_per_subject=19,
                mem_ecc=0, load_balancer=1, load_factor=4, subject_chunk=200, base_value=100)
    assert hb.runtime_options() == want
    assert O.runtime_options() == want


def test_options_parse_by_hand():
    want = dict(gpu_ratio=float(np.float32(0.7)), fixed_ratio=0, capability_rate=10.5,
                max_hits_per_subject=30, mem_ecc=1, load_balancer=0, load_factor=6,
                subject_chunk=150, base_value=80)
    assert hb.runtime_options(OPTS_TEXT) == want
    assert O.runtime_options(OPTS_TEXT) == want


def test_options_file_the_reference_ships():
    # tests/golden/H-BLAST_runtime_options is the reference's own options file
    # (runtime_setting_files/H-BLAST_runtime_options), kept as data: ratio 1,
    # adaptive, capability 12, 20 HSPs per subject, no ECC ("0 " with a
    # trailing blank), new balancer, factor 4, chunk 200, base 100
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "H-BLAST_runtime_options")
    with open(path, "rb") as f:
        text = f.read()
    want = dict(gpu_ratio=1.0, fixed_ratio=0, capability_rate=12.0, max_hits_per_subject=20,
                mem_ecc=0, load_balancer=1, load_factor=4, subject_chunk=200, base_value=100)
    assert hb.runtime_options(text) == want
    assert O.runtime_options(text) == want


def test_options_fallbacks_by_hand():
    # out-of-range values take the reference's fallbacks; "2" is not "1" for
    # the fixed flag; atoi/atof read a number's leading part
    text = ("#\n1.5\n#\n2\n#\n-3\n#\n0\n#\n7\n#\n-1\n#\n-2\n#\n-5\n#\n-9\n")
    o = hb.runtime_options(text)
    assert o == O.runtime_options(text)
    assert o == dict(gpu_ratio=0.0, fixed_ratio=0, capability_rate=12.0,
                     max_hits_per_subject=20, mem_ecc=1, load_balancer=1, load_factor=4,
                     subject_chunk=200, base_value=100)
    text = "#\n 0.25x\n#\n1 \n#\n3e1\n#\n+8\n#\n0\n#\n00\n#\n12abc\n#\n\n#\n0x10\n"
    o = hb.runtime_options(text)
    assert o == O.runtime_options(text)
    assert (o["gpu_ratio"], o["fixed_ratio"], o["capability_rate"], o["max_hits_per_subject"],
            o["load_balancer"], o["load_factor"], o["subject_chunk"], o["base_value"]) == \
        (0.25, 1, 30.0, 8, 0, 12, 0, 0)


def test_options_short_file_is_an_error():
    text = OPTS_TEXT.rsplit("\n", 2)[0] + "\n"     # 17 lines
    assert O.runtime_options(text) is None
    with pytest.raises(hb.HBError) as e:
        hb.runtime_options(text)
    assert e.value.code == -9
    # no trailing newline after the last value is fine
    assert hb.runtime_options(OPTS_TEXT[:-1]) == hb.runtime_options(OPTS_TEXT)


def test_options_random_texts_match_oracle():
    rng = np.random.default_rng(5)
    tokens = ["0", "1", "-1", "0.5", "1e-3", "-0.0", "12", "  7", "x", "", "3.9", "2147483648",
              "-2147483649", "99999999999", "nan", "inf", "-inf", "0.999999", "1.0000001"]
    for _ in range(2000):
        vals = [tokens[rng.integers(len(tokens))] for _ in range(9)]
        text = "".join(f"#c{i}\n{v}\n" for i, v in enumerate(vals))
        a, b = hb.runtime_options(text), O.runtime_options(text)
        for k in a:
            assert (a[k] == b[k]) or (np.isnan(a[k]) and np.isnan(b[k])), (text, k)


# --- split (functions.cpp:449-571) --------------------------------------------

def test_split_by_hand():
    # grid 2 x 3 (groups of 6), group lengths 16, 32, 48, 16 sequences (the
    # last group holds 4, two blocks of 3 slots): workload 16*6 + 32*6 +
    # 48*4 = 480; the file is 96 + 192 + 288 = 576 bytes.
    gl, n, nb = [16, 32, 48], 16, 576
    for split in (hb.hdb_split, O.split):
        # ratio 0.5 -> target 240: without the last group 288 > 240, without
        # group 1 too 96 <= 240; rest 144 fills trunc(144 / (32*3) * 3) = 4
        # sequences of group 1 (4.5 truncated) -> 10 sequences, 96 + 32 * 6
        # bytes (two blocks), workload 96 + 128 = 224
        share, r, f = split(gl, n, nb, 2, 3, 4, 0.5, 1)
        assert share == dict(gpu_seqs=10, n_groups=2, bytes=288) and (r, f) == (0.5, 1)
        share, r, f = split(gl, n, nb, 2, 3, 4, 0.5, 0)
        assert share["gpu_seqs"] == 10 and f == 0
        assert r == float(np.float32(224 / 480))
        # ratio 1 (and above): the whole .hDb
        for ratio in (1.0, 1.5):
            assert split(gl, n, nb, 2, 3, 4, ratio, 0)[0] == dict(gpu_seqs=16, n_groups=3,
                                                                  bytes=576)
        # ratio 0: nothing; the realised ratio 0
        share, r, f = split(gl, n, nb, 2, 3, 4, 0.0, 0)
        assert share == dict(gpu_seqs=0, n_groups=0, bytes=0) and r == 0.0
        # ratio 0.25 -> target 120: whole group 0 (96) + trunc(24 / 96 * 3) = 0
        share, r, f = split(gl, n, nb, 2, 3, 4, 0.25, 0)
        assert share == dict(gpu_seqs=6, n_groups=1, bytes=96) and r == float(np.float32(0.2))
        # ratio 0.9 (a float: 0.89999998) -> target 431.99999: without the
        # last group 288 <= target, rest 143.99999 fills trunc(0.9999999 * 3)
        # = 2 of the last group's 4 sequences (3 with an exact 0.9)
        share, r, f = split(gl, n, nb, 2, 3, 4, 0.9, 0)
        assert share == dict(gpu_seqs=14, n_groups=3, bytes=96 + 192 + 144)
        assert r == float(np.float32(384 / 480))
        # ratio 0.95 -> target 456: rest 168 fills trunc(168 / 144 * 3) = 3
        share, r, f = split(gl, n, nb, 2, 3, 4, 0.95, 0)
        assert share == dict(gpu_seqs=15, n_groups=3, bytes=96 + 192 + 144)
        assert r == float(np.float32(432 / 480))


def test_split_start_ratio_by_thread_count():
    # -1 = no ratio set: 1 thread -> 1 (nothing cut), 2 -> 0.9 (fixed stays,
    # the ratio stays -1), 3 and up -> 0.9 ... 0.4, adaptive from then on
    gl = [16] * 10
    n, nb = 60, 16 * 60
    for split in (hb.hdb_split, O.split):
        share, r, f = split(gl, n, nb, 2, 3, 1)
        assert share["gpu_seqs"] == 60 and (r, f) == (-1.0, 1)
        share, r, f = split(gl, n, nb, 2, 3, 2)
        assert share["gpu_seqs"] == 54 and (r, f) == (-1.0, 1)
        for threads, ratio in [(3, 0.9), (4, 0.8), (5, 0.75), (6, 0.65), (7, 0.55), (8, 0.45),
                               (9, 0.4), (64, 0.4), (0, 0.4)]:
            share, r, f = split(gl, n, nb, 2, 3, threads)
            assert f == 0
            # equal group lengths: the share is trunc(ratio * 60) sequences
            # up to the double rounding of the fill
            assert abs(share["gpu_seqs"] - ratio * 60) < 1, threads
            assert r == float(np.float32(share["gpu_seqs"] / 60))


def brute_split(gl, n, nb, blocksx, threadsx, ratio, fixed):
    """The split restated independently: the whole groups kept are the
    longest prefix whose workload is within the target."""
    G = blocksx * threadsx
    ng = len(gl)
    cost = [float(gl[g]) * (G if g < ng - 1 else n - (ng - 1) * G) for g in range(ng)]
    total = sum(cost[:-1]) + cost[-1] if ng else 0.0
    r = float(ratio)
    target = total * r
    if not ng or not r < 1:
        return dict(gpu_seqs=n, n_groups=ng, bytes=nb), ratio
    prefix = [0.0]
    for g in range(ng - 1):
        prefix.append(prefix[-1] + float(gl[g]) * G)
    k = max(j for j in range(ng) if prefix[j] <= target or j == 0)
    kept = prefix[k]
    L = gl[k]
    added = int(((target - kept) / (L * threadsx)) * threadsx) if L else 0
    kept += float(L) * added
    with np.errstate(invalid="ignore"):
        q = np.float32(np.float64(kept) / np.float64(total)) if total else np.float32("nan")
    new_ratio = ratio if fixed == 1 else (float(q) if q != q else float(min(max(q, 0), 1)))
    if k * G + added < n:
        b = sum(int(x) * G for x in gl[:k]) + (L * (-(-added // threadsx)) * threadsx if added
                                               else 0)
        return dict(gpu_seqs=k * G + added, n_groups=k + (1 if added else 0), bytes=b), new_ratio
    return dict(gpu_seqs=n, n_groups=ng, bytes=nb), new_ratio


def same_float(a, b):
    return a == b or (a != a and b != b)


def test_split_random_matches_brute_force_and_oracle():
    rng = np.random.default_rng(11)
    for it in range(3000):
        blocksx, threadsx = int(rng.integers(1, 6)), int(rng.integers(1, 70))
        G = blocksx * threadsx
        ng = int(rng.integers(1, 12))
        n = (ng - 1) * G + int(rng.integers(1, G + 1))
        gl = np.sort(rng.integers(0, 40, ng) * 16).astype(np.int32)
        last_slots = -(-(n - (ng - 1) * G) // threadsx) * threadsx
        nb = int(gl[:-1].sum()) * G + int(gl[-1]) * last_slots
        ratio = float(np.float32(rng.choice([rng.random(), 0.0, 1.0, rng.random() * 2])))
        fixed = int(rng.integers(0, 2))
        got = hb.hdb_split(gl, n, nb, blocksx, threadsx, 4, ratio, fixed)
        ref = O.split(gl, n, nb, blocksx, threadsx, 4, ratio, fixed)
        want, want_r = brute_split([int(x) for x in gl], n, nb, blocksx, threadsx, ratio, fixed)
        assert got[0] == ref[0] == want and got[2] == ref[2], (it, gl, n, blocksx, threadsx, ratio)
        # all-empty groups weigh nothing: the realised ratio is 0 / 0 (NaN) in all three
        assert same_float(got[1], ref[1]) and same_float(got[1], want_r), it
        if got[1] == got[1] and fixed == 0 and ratio < 1:
            # the realised ratio never exceeds the asked-for one by more than
            # float rounding, and the share is a prefix ending in whole blocks
            assert got[1] <= ratio + 1e-6
        assert got[0]["gpu_seqs"] <= n and got[0]["bytes"] <= nb


def test_split_share_is_a_loadable_prefix_of_the_file():
    # the .hDb bytes up to the share's size, read with the share's counts and
    # group lengths, hold exactly the volume's first gpu_seqs sequences
    rng = np.random.default_rng(3)
    lens = np.sort(rng.integers(1, 300, 5000)).astype(np.uint64)
    offs = np.zeros(lens.size + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    res = rng.integers(1, 25, int(offs[-1]), dtype=np.uint8)
    n = lens.size
    for blocksx, threadsx in [(4, 64), (3, 50), (1, 7)]:
        data, gl = O.hdb_pack(res, offs, n, blocksx, threadsx)
        for ratio in (0.05, 0.31, 0.5, 0.77, 0.99):
            share, _, _ = hb.hdb_split(gl, n, data.size, blocksx, threadsx, 4, ratio, 1)
            k = share["gpu_seqs"]
            assert 0 < k < n
            r2, o2 = O.hdb_unpack(data[:share["bytes"]], k, blocksx, threadsx,
                                  gl[:share["n_groups"]])
            assert np.array_equal(o2, offs[:k + 1])
            assert np.array_equal(r2, res[:int(offs[k])])


def test_split_rejects_bad_arguments():
    for args in [([16, 16], 5, 0, 2, 3, 4, 0.5, 1),    # 5 sequences are one group of 6
                 ([16], 7, 0, 2, 3, 4, 0.5, 1),        # 7 sequences need two groups
                 ([16], 5, 0, 0, 3, 4, 0.5, 1),        # blocksx 0
                 ([16], 5, 0, 2, 3, 4, -0.5, 1),       # negative ratio other than -1
                 ([-16], 5, 0, 2, 3, 4, 0.5, 1)]:
        with pytest.raises(hb.HBError):
            hb.hdb_split(*args)
    # an empty .hDb keeps nothing to cut
    assert hb.hdb_split([], 0, 0, 2, 3, 4, 0.5, 0)[0] == dict(gpu_seqs=0, n_groups=0, bytes=0)


# --- next ratio (functions.cpp:51-140) --------------------------------------

def opts(**kw):
    o = hb.runtime_options()
    o.update(kw)
    return o


def both(o, *a, **kw):
    x = hb.balance_next_ratio(o, *a, **kw)
    y = O.next_ratio(o, *a, **kw)
    assert x == y, (a, kw, x, y)
    return x


def test_next_ratio_by_hand():
    # equal volumes, no HSP charge, 12 CPU threads and one GPU at capability
    # 12: the GPU counts as 2 cards in all; M = 0.65 / (0.65 + 0.35) = 0.65.
    o = opts(load_balancer=0)
    a = (1, 0.5, 12, [100, 100, 100], [0.5, 0.5], [0, 0])
    # the CPU took twice the GPU's time: +1 x 0.65 / 2 = 0.325 of a card,
    # / 0.96 -> 0.5 + 0.33854 (the step is a float: 0.325 rounds first)
    r = both(o, *a, 1.0, 2.0, query_length=500)
    assert r == float(np.float32(0.5 + float(np.float32(0.325)) / 0.96))
    # the GPU took twice the CPU's time: -0.5 x (1 - 0.325) = -0.3375
    r = both(o, *a, 2.0, 1.0, query_length=500)
    assert r == float(np.float32(0.5 - float(np.float32(0.3375)) / 0.96))
    # equal times: unchanged
    assert both(o, *a, 1.5, 1.5, query_length=500) == 0.5
    # a GPU time of 0 gives 0.9
    assert both(o, *a, 0.0, 3.0) == float(np.float32(0.9))
    # far apart: clamped to 1 and to 0
    assert both(o, *a, 1.0, 100.0) == 1.0
    assert both(o, *a, 100.0, 1.0) == 0.0


def test_next_ratio_volume_sizes_scale_the_times():
    # the CPU time is scaled to the CPU's next share (bytes x (1 - 0.96 r)):
    # volume 2 twice volume 1 -> the CPU's 1 s counts as 2 s; the GPU time to
    # the next volume's size: volume 3 (ring slot 0) the size of volume 2
    o = opts(load_balancer=0)
    r = both(o, 2, 0.5, 12, [200, 100, 200], [0.5, 0.5], [0, 0], 2.0, 1.0, query_length=500)
    assert r == 0.5
    # long queries damp the GPU's growth: a next volume (ring slot 2) twice
    # this one (slot 1) counts as 2, 2^0.9, 2^0.8, 2^0.75
    for ql, e in [(1000, 1.0), (2000, 0.9), (5000, 0.8), (9000, 0.75)]:
        gpu = 1.0 / (2.0 ** e)   # scaled, equal to the CPU's 1 s
        r = both(o, 1, 0.5, 12, [100, 100, 200], [0.5, 0.5], [0, 0], gpu, 1.0, query_length=ql)
        assert abs(r - 0.5) < 1e-6, (ql, r)


def test_next_ratio_hsp_charge_keeps_the_reference_divisor():
    # ungapped with the balancer on: each GPU HSP adds load_factor / (0.96 r
    # + 0.03) to this volume's CPU load, but the previous volume's charge is
    # divided by the just-computed load of this volume (functions.cpp:93).
    # bytes 1000 each, r 0.5 -> 0.48, 100 HSPs each, factor 4:
    #   now    = 520 + 400 / 0.51          = 1304.3137...
    #   before = 520 + 400 / (now + 0.03)  = 520.3066...
    o = opts(load_balancer=1, load_factor=4)
    now = 1000 * (1 - 0.48) + 100 * 4 / (0.48 + 0.03)
    before = 1000 * (1 - 0.48) + 100 * 4 / (now + 0.03)
    scale = now / before
    assert 2.5 < scale < 2.51
    r = both(o, 1, 0.5, 12, [1000, 1000, 1000], [0.5, 0.5], [100, 100], 1.0, 0.5,
             query_length=500)
    step = float(np.float32(float(np.float32(0.5 * scale - 1.0)) * 0.325))
    assert abs(r - (0.5 + step / 0.96)) < 1e-6
    # with the charge divided by its own rate the scale would be 1 and the
    # ratio would fall instead
    assert r > 0.58


def test_next_ratio_gapped_charge_by_query_length():
    # gapped: per HSP base + (750 - base) x min(1500, avg query length) /
    # 1500, over (0.96 r + 0.03), in both volumes.  Volume 1 (ring slot 1)
    # has 2000 residues and 50 HSPs, volume 0 1000 and 10; the next (slot 2)
    # is half of volume 1, so the GPU's 1 s counts as 0.5 s.
    o = opts(base_value=100)
    for avg, per in [(750, 425.0), (3000, 750.0), (0, 100.0)]:
        now = 2000 * 0.52 + 50 * per / 0.51
        before = 1000 * 0.52 + 10 * per / 0.51
        r = both(o, 1, 0.5, 12, [1000, 2000, 1000], [0.5, 0.5], [10, 50], 1.0, 0.1,
                 gapped=True, query_length=500, avg_query_length=avg)
        step = float(np.float32((0.1 * now / before - 0.5) / 0.5))
        step = float(np.float32(step * (0.325 if step > 0 else 0.675)))
        assert 0 < r < 1
        assert abs(r - (0.5 + step / 0.96)) < 1e-6, avg


def test_next_ratio_errors_where_the_reference_exits():
    o = opts()
    assert O.next_ratio(o, 1, 0.5, 12, [0, 1, 1], [0.5, 0.5], [0, 0], 1.0, 1.0) is None
    assert O.next_ratio(o, 1, 0.5, 12, [1, 1, 1], [0.0, 0.5], [0, 0], 1.0, 1.0) is None
    for args in [(1, 0.5, 12, [0, 1, 1], [0.5, 0.5], [0, 0]),
                 (1, 0.5, 12, [1, 1, 1], [0.5, 0.0], [0, 0]),
                 (0, 0.5, 12, [1, 1, 1], [0.5, 0.5], [0, 0])]:
        with pytest.raises(hb.HBError):
            hb.balance_next_ratio(o, *args, 1.0, 1.0)


def test_next_ratio_random_matches_oracle():
    rng = np.random.default_rng(17)
    for it in range(20000):
        o = opts(load_balancer=int(rng.integers(0, 2)), load_factor=int(rng.integers(0, 9)),
                 base_value=int(rng.integers(0, 800)),
                 capability_rate=float(rng.choice([12.0, rng.random() * 30 + 0.1])))
        vb = [int(x) for x in rng.integers(1, 10 ** 10, 3)]
        rr = [float(np.float32(x)) for x in rng.random(2) * 1.04 + 1e-3]
        hr = [int(x) for x in rng.integers(0, 10 ** 6, 2)]
        gpu = float(rng.choice([0.0, rng.random() * 10]))
        cpu = float(rng.random() * 10)
        r = float(np.float32(rng.random()))
        args = (int(rng.integers(1, 50)), r, int(rng.integers(1, 129)), vb, rr, hr, gpu, cpu)
        kw = dict(gapped=bool(rng.integers(0, 2)), query_length=int(rng.integers(0, 12000)),
                  avg_query_length=int(rng.integers(0, 4000)),
                  gpu_cards=int(rng.integers(1, 9)))
        x = both(o, *args, **kw)
        assert 0.0 <= x <= 1.0
