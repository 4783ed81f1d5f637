"""H-BLAST GPU database volumes (.hDb) and index (.hDbI): oracle pins and the
product's host-side writer and index parser (no GPU).

The oracle's writer restates CreateH_BLAST_database
(ncbi_blast_files/functions.cpp:785-1165) and its reader the GPU kernel's
subject walk (H-BLAST/h_blast_kernel.cu:593-608,1060-1064); they are pinned
by hand-worked byte layouts, hand-computed GPU shares and by agreeing with
each other (two independent pieces of the reference).
"""
import os

import numpy as np
import pytest

import oracle_lib as O

hb = pytest.importorskip("hbseed")


def sorted_db(seed, n, lo=0, hi=300, zero_free=True):
    rng = np.random.default_rng(seed)
    lens = np.sort(rng.integers(lo, hi + 1, n)).astype(np.uint64)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    res = rng.integers(1 if zero_free else 0, 25, int(offs[-1]), dtype=np.uint8)
    return res, offs


# --- GPU share of a volume (functions.cpp:842-843) -------------------------

@pytest.mark.parametrize("n,expect", [
    (0, 0), (20, 0), (32, 0),        # 31.968 -> 31 -> 0
    (33, 32), (1000, 992),           # 999 -> 992
    (1001, 992),                     # 999.999 -> 999 -> 992
    (3000, 2976),                    # 2997 -> 2976
    (10 ** 6, 998976),               # 999000 -> 998976
    (2_000_000, 1_997_984),          # 1998000 -> 1997984
])
def test_gpu_share_hand_values(n, expect):
    assert O.hdb_gpu_seqs(n) == expect
    assert hb.hdb_gpu_seqs(n) == expect


def test_gpu_share_product_matches_oracle():
    ns = list(range(0, 5000)) + list(range(10 ** 6, 10 ** 6 + 3000, 7)) + [2 ** 31 - 1]
    assert [hb.hdb_gpu_seqs(n) for n in ns] == [O.hdb_gpu_seqs(n) for n in ns]


# --- byte layout, worked by hand --------------------------------------------

def test_layout_one_block_by_hand():
    # blocksx 1, threadsx 2: sequences AB (1,2) and CDEFG (3..7); the group
    # length is 5 rounded up to 16, so 2 x 16 bytes.  Row 0 holds residues
    # 0-3 of sequence 0 then of sequence 1, row 1 residues 4-7.
    res = np.array([1, 2, 3, 4, 5, 6, 7], np.uint8)
    offs = np.array([0, 2, 7], np.uint64)
    expect = np.zeros(32, np.uint8)
    expect[0:2] = [1, 2]
    expect[4:8] = [3, 4, 5, 6]
    expect[12] = 7
    for pack in (O.hdb_pack, hb.hdb_pack):
        data, gl = pack(res, offs, 2, 1, 2)
        assert gl.tolist() == [16]
        assert data.tolist() == expect.tolist()
    r2, o2 = O.hdb_unpack(expect, 2, 1, 2, [16])
    assert r2.tolist() == res.tolist() and o2.tolist() == offs.tolist()


def test_layout_two_blocks_by_hand():
    # blocksx 2, threadsx 2, lengths 1,1,2,3: one group of 4 (length 16);
    # block 1 (sequences 2, 3) starts at 1 * 2 * 16 = 32.
    res = np.array([9, 8, 7, 6, 5, 4, 3], np.uint8)
    offs = np.array([0, 1, 2, 4, 7], np.uint64)
    expect = np.zeros(64, np.uint8)
    expect[0] = 9
    expect[4] = 8
    expect[32:34] = [7, 6]
    expect[36:39] = [5, 4, 3]
    for pack in (O.hdb_pack, hb.hdb_pack):
        data, gl = pack(res, offs, 4, 2, 2)
        assert gl.tolist() == [16]
        assert data.tolist() == expect.tolist()


def test_layout_partial_last_group_by_hand():
    # blocksx 1, threadsx 2, lengths 1, 2, 17: group 0 = sequences 0,1
    # (length 16, 32 bytes); group 1 = sequence 2 alone, length 32, still a
    # whole block of 2 slots (64 bytes) -> 96 bytes.  Sequence 2's stripes
    # sit every 2 words from byte 32.
    s2 = np.arange(1, 18, dtype=np.uint8)
    res = np.concatenate([[20], [21, 22], s2]).astype(np.uint8)
    offs = np.array([0, 1, 3, 20], np.uint64)
    expect = np.zeros(96, np.uint8)
    expect[0] = 20
    expect[4:6] = [21, 22]
    expect[32:36] = s2[0:4]
    expect[40:44] = s2[4:8]
    expect[48:52] = s2[8:12]
    expect[56:60] = s2[12:16]
    expect[64] = s2[16]
    for pack in (O.hdb_pack, hb.hdb_pack):
        data, gl = pack(res, offs, 3, 1, 2)
        assert gl.tolist() == [16, 32]
        assert data.tolist() == expect.tolist()


def test_layout_sizes_by_hand():
    # 5 sequences of length 10 in groups of 2 x 3 = 6: one partial group of
    # 5 members -> ceil(5 / 3) = 2 blocks x 3 slots x 16 = 96 bytes
    offs = np.arange(0, 60, 10, dtype=np.uint64)
    assert O.hdb_layout(offs, 5, 2, 3)[1] == 96
    assert hb.hdb_layout(offs, 5, 2, 3)[1] == 96
    # 7 sequences: a full group (6 x 16) + 1 block of 3 x 16
    offs = np.arange(0, 80, 10, dtype=np.uint64)
    assert O.hdb_layout(offs, 7, 2, 3)[1] == 96 + 48
    assert hb.hdb_layout(offs, 7, 2, 3)[1] == 96 + 48


# --- writer vs reader, product vs oracle ------------------------------------

GRIDS = [(1, 1), (1, 2), (2, 3), (3, 64), (2, 192), (512, 192), (7, 100)]


@pytest.mark.parametrize("blocksx,threadsx", GRIDS)
def test_writer_matches_oracle_and_reader(blocksx, threadsx):
    res, offs = sorted_db(blocksx * 1000 + threadsx, 1500, 0, 400)
    n = len(offs) - 1
    d_o, gl_o = O.hdb_pack(res, offs, n, blocksx, threadsx)
    d_p, gl_p = hb.hdb_pack(res, offs, n, blocksx, threadsx)
    assert gl_p.tolist() == gl_o.tolist()
    assert np.array_equal(d_p, d_o)
    r2, o2 = O.hdb_unpack(d_o, n, blocksx, threadsx, gl_o)
    assert np.array_equal(o2, offs) and np.array_equal(r2, res)


def test_trailing_gap_code_is_not_recovered():
    # the file keeps no lengths: a sequence ends after its last non-zero byte
    res = np.array([1, 2, 0, 3, 0, 0], np.uint8)
    offs = np.array([0, 3, 6], np.uint64)
    d, gl = O.hdb_pack(res, offs, 2, 1, 2)
    r2, o2 = O.hdb_unpack(d, 2, 1, 2, gl)
    assert o2.tolist() == [0, 2, 3]
    assert r2.tolist() == [1, 2, 3]


def test_writer_rejects_unsorted_and_bad_grid():
    offs = np.array([0, 5, 7], np.uint64)
    with pytest.raises(hb.HBError):
        hb.hdb_layout(offs, 2, 1, 2)
    offs = np.array([0, 5, 10], np.uint64)
    for bx, tx in [(0, 2), (1, 0), (65537, 1), (1, 65537)]:
        with pytest.raises(hb.HBError):
            hb.hdb_layout(offs, 2, bx, tx)


# --- index text --------------------------------------------------------------

def volumes_of(seed, sizes, blocksx, threadsx):
    vols, first = [], 0
    for k, vs in enumerate(sizes):
        res, offs = sorted_db(seed + k, vs, 1, 200)
        gs = O.hdb_gpu_seqs(vs)
        data, gl = O.hdb_pack(res, offs, gs, blocksx, threadsx)
        vols.append(dict(name=f"db.{k:02d}.hDb", bytes=int(data.size), volume_seqs=vs,
                         gpu_seqs=gs, store_limit=first + gs, break_limit=first + vs,
                         group_len=gl, cpu_rest_bytes=int(offs[vs] - offs[gs])))
        first += vs
    return vols


def test_index_text_by_hand():
    v = dict(name="nr.hDb", bytes=96, volume_seqs=40, gpu_seqs=32, store_limit=32,
             break_limit=40, group_len=[16, 32], cpu_rest_bytes=123)
    expect = ("#Number of blocksx\n2\n#Number of threadsx\n3\n#Percentage\n99\n"
              "#Number of volumes\n1\n#Volume name\nnr.hDb\n#Volume size in bytes\n96\n"
              "#Volume sequences\n40\n#Number of sequences in GPU volume\n32\n"
              "#Store limit\n32\n#Break limit\n40\n#Group number\n2\n"
              "#Length of each group\n16\n32\n"
              "#Length of the rest sequences in CPU database\n123\n")
    assert O.hdbi_format(2, 3, [v]) == expect
    assert hb.hdbi_format(2, 3, [v]) == expect


def test_index_product_matches_oracle_and_round_trips():
    vols = volumes_of(7, [5000, 777, 33, 20], 4, 64)
    text = hb.hdbi_format(4, 64, vols)
    assert text == O.hdbi_format(4, 64, vols)
    idx = hb.hdbi_parse(text)
    assert (idx["blocksx"], idx["threadsx"], idx["percentage"]) == (4, 64, 99)
    assert len(idx["volumes"]) == len(vols)
    for a, b in zip(idx["volumes"], vols):
        for k in ("name", "bytes", "volume_seqs", "gpu_seqs", "store_limit", "break_limit",
                  "cpu_rest_bytes"):
            assert a[k] == b[k], k
        assert a["group_len"].tolist() == list(b["group_len"])
    # CRLF line ends and a trailing blank line are accepted
    idx2 = hb.hdbi_parse(text.replace("\n", "\r\n") + "\n")
    assert [v["group_len"].tolist() for v in idx2["volumes"]] == \
        [v["group_len"].tolist() for v in idx["volumes"]]


def error_offset(text):
    with pytest.raises(hb.HBError) as e:
        hb.hdbi_parse(text)
    assert e.value.code == -9
    return int(str(e.value).split("text offset ")[1].split(")")[0])


def test_index_errors_name_the_line():
    v = volumes_of(3, [100], 1, 2)
    text = hb.hdbi_format(1, 2, v)
    lines = text.split("\n")
    starts = np.cumsum([0] + [len(x) + 1 for x in lines])
    # a value that is not a number
    bad = lines.copy()
    bad[5] = "x7"
    assert error_offset("\n".join(bad)) == starts[5]
    # a missing comment line (the value lands where a comment belongs)
    bad = lines[:2] + lines[3:]
    assert error_offset("\n".join(bad)) == starts[2]
    # truncated inside the group lengths
    k = lines.index("#Length of each group")
    assert error_offset("\n".join(lines[:k + 1])) >= starts[k]
    # text after the last volume
    assert error_offset(text + "junk\n") == len(text)
    # grid out of range
    bad = lines.copy()
    bad[1] = "0"
    assert error_offset("\n".join(bad)) == starts[1]


def test_write_hdb_files_read_back(tmp_path):
    res, offs = sorted_db(11, 3000, 1, 300)
    # two volumes, each sorted on its own
    n0 = 1800
    prefix = str(tmp_path / "db")
    idx = hb.write_hdb(prefix, res, offs, volume_seqs=[n0, 3000 - n0], blocksx=3,
                       threadsx=32)
    assert [v["name"] for v in idx["volumes"]] == ["db.00.hDb", "db.01.hDb"]
    with open(prefix + ".hDbI") as f:
        assert hb.hdbi_parse(f.read())["volumes"][1]["store_limit"] == \
            n0 + O.hdb_gpu_seqs(3000 - n0)
    first = 0
    for v in idx["volumes"]:
        data = np.fromfile(os.path.join(tmp_path, v["name"]), np.uint8)
        assert data.size == v["bytes"]
        r2, o2 = O.hdb_unpack(data, v["gpu_seqs"], 3, 32, v["group_len"])
        lo = int(offs[first])
        want_o = offs[first:first + v["gpu_seqs"] + 1] - offs[first]
        assert np.array_equal(o2, want_o)
        assert np.array_equal(r2, res[lo:lo + int(want_o[-1])])
        rest = int(offs[first + v["volume_seqs"]] - offs[first + v["gpu_seqs"]])
        assert v["cpu_rest_bytes"] == rest
        first += v["volume_seqs"]


# --- sorted volumes (makeblastdb -sort_volumes) ------------------------------

def offs_of(lens):
    o = np.zeros(len(lens) + 1, np.uint64)
    o[1:] = np.cumsum(np.asarray(lens, np.uint64))
    return o


def test_sort_volumes_by_hand():
    # sequence file: offset 1, +len+1 per sequence; a later sequence fits
    # while offset + len < max (writedb_files.hpp:435-494)
    # max 40: 1 -> 12 (len 10) -> 33 (len 20, 12 + 20 < 40); len 30: 33 + 30
    # >= 40 -> new volume: 1 -> 32; len 5: 32 + 5 < 40 -> 38
    order, vs = hb.sort_volumes(offs_of([10, 20, 30, 5]), 40)
    assert vs.tolist() == [2, 2]
    assert order.tolist() == [0, 1, 3, 2]        # each volume shortest first
    # offset + len == max does not fit (strict)
    order, vs = hb.sort_volumes(offs_of([10, 20]), 32)
    assert vs.tolist() == [1, 1]
    order, vs = hb.sort_volumes(offs_of([10, 20]), 33)
    assert vs.tolist() == [2]
    # a volume's first sequence fits whatever its length
    # 100 opens volume 0 (offset 102); 3: 102 + 3 >= 50 opens volume 1
    # (offset 5); 200: 5 + 200 >= 50 opens volume 2
    order, vs = hb.sort_volumes(offs_of([100, 3, 200]), 50)
    assert vs.tolist() == [1, 1, 1]
    assert order.tolist() == [0, 1, 2]
    # the letter limit: letters + len > limit cuts
    order, vs = hb.sort_volumes(offs_of([4, 4, 4]), 10 ** 6, max_letters=8)
    assert vs.tolist() == [2, 1]
    # descending input comes out ascending inside the volume
    order, vs = hb.sort_volumes(offs_of([9, 7, 5, 3]), 10 ** 6)
    assert vs.tolist() == [4] and order.tolist() == [3, 2, 1, 0]


def cut_restated(lens, max_bytes, max_letters=0):
    """The volume cut written as the file sizes after each sequence: a new
    volume starts at the first sequence (after a volume's first) whose
    sequence-file size would reach max_bytes or whose letters would pass
    max_letters."""
    out, start = [], 0
    for i in range(1, len(lens) + 1):
        if i == len(lens):
            out.append(i - start)
            break
        size = 1 + sum(int(x) + 1 for x in lens[start:i])   # file size with start..i-1
        letters = sum(int(x) for x in lens[start:i])
        if size + int(lens[i]) >= max_bytes or (max_letters and letters + int(lens[i]) > max_letters):
            out.append(i - start)
            start = i
    return out


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sort_volumes_random(seed):
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 400, 3000)
    lens[rng.random(3000) < 0.3] = 77                 # many ties
    offs = offs_of(lens)
    for max_bytes, max_letters in [(20_000, 0), (100_000, 30_000), (10 ** 9, 0), (300, 0)]:
        order, vs = hb.sort_volumes(offs, max_bytes, max_letters)
        assert vs.tolist() == cut_restated(lens, max_bytes, max_letters)
        first = 0
        for c in vs:
            got = order[first:first + c]
            assert sorted(got.tolist()) == list(range(first, first + c))   # the volume's own
            assert np.all(np.diff(lens[got]) >= 0)                          # shortest first
            first += c


def test_build_hdb_from_unsorted_sequences(tmp_path):
    rng = np.random.default_rng(9)
    lens = rng.integers(1, 300, 2500).astype(np.uint64)
    offs = offs_of(lens)
    res = rng.integers(1, 25, int(offs[-1]), dtype=np.uint8)
    prefix = str(tmp_path / "built")
    idx, order = hb.build_hdb(prefix, res, offs, max_file_bytes=150_000, blocksx=2, threadsx=32)
    vols = idx["volumes"]
    assert len(vols) > 1
    assert [v["name"] for v in vols] == [f"built.{k:02d}.hDb" for k in range(len(vols))]
    first = 0
    for v in vols:
        data = np.fromfile(os.path.join(tmp_path, v["name"]), np.uint8)
        assert data.size == v["bytes"]
        r2, o2 = O.hdb_unpack(data, v["gpu_seqs"], 2, 32, v["group_len"])
        ids = order[first:first + v["gpu_seqs"]]
        want = np.concatenate([res[int(offs[i]):int(offs[i + 1])] for i in ids]) \
            if len(ids) else np.zeros(0, np.uint8)
        assert np.array_equal(o2[1:] - o2[:-1], lens[ids])
        assert np.array_equal(r2, want)
        rest = order[first + v["gpu_seqs"]:first + v["volume_seqs"]]
        assert v["cpu_rest_bytes"] == int(lens[rest].sum())
        assert v["break_limit"] == first + v["volume_seqs"]
        first += v["volume_seqs"]
    assert first == 2500
