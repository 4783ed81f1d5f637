"""GPU parity of the .hDb volume loader (hb_db_upload_hdb): the database the
device unpacks equals what the oracle's restatement of the reference
kernel's subject walk reads from the same file, byte for byte; scans of a
loaded volume equal scans of the same sequences uploaded as packed arrays.
"""
import sys

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


def sorted_db(seed, n, lo=1, hi=400, zero_rate=0.0):
    rng = np.random.default_rng(seed)
    lens = np.sort(rng.integers(lo, hi + 1, n)).astype(np.uint64)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    res = rng.integers(1, 25, int(offs[-1]), dtype=np.uint8)
    if zero_rate:
        res[rng.random(res.size) < zero_rate] = 0
    return res, offs


def load_and_compare(ctx, res, offs, blocksx, threadsx):
    n = len(offs) - 1
    data, gl = hbseed.hdb_pack(res, offs, n, blocksx, threadsx)
    print(f"hdb load n={n} grid={blocksx}x{threadsx} bytes={data.size}", flush=True)
    db = hbseed.Database.from_hdb(ctx, data, blocksx, threadsx, n, gl)
    try:
        got_r, got_o = db.fetch()
    finally:
        db.close()
    exp_r, exp_o = oracle_lib.hdb_unpack(data, n, blocksx, threadsx, gl)
    assert np.array_equal(got_o, exp_o)
    assert np.array_equal(got_r, exp_r)
    return got_r, got_o


@pytest.mark.parametrize("blocksx,threadsx,n,hi", [
    (1, 1, 300, 200), (1, 2, 301, 50), (2, 3, 1000, 400), (3, 64, 5000, 600),
    (2, 100, 4321, 900), (512, 192, 20000, 800), (4, 192, 50000, 2000), (1, 65, 130, 5000),
])
def test_loader_matches_oracle_reader(ctx, blocksx, threadsx, n, hi):
    res, offs = sorted_db(blocksx * 7919 + threadsx, n, 1, hi)
    got_r, got_o = load_and_compare(ctx, res, offs, blocksx, threadsx)
    # sequences without a trailing gap code come back unchanged
    assert np.array_equal(got_o, offs) and np.array_equal(got_r, res)


def test_loader_gap_codes_and_empty_sequences(ctx):
    # interior and trailing code-0 residues, empty sequences
    res, offs = sorted_db(5, 3000, 0, 120, zero_rate=0.05)
    got_r, got_o = load_and_compare(ctx, res, offs, 3, 50)
    assert got_o[-1] <= offs[-1]


def test_scan_of_loaded_volume_equals_packed_upload(ctx):
    res, offs = hbseed.make_db(2100, 30000)
    s_res, s_offs, order = hbseed.sort_by_length(res, offs)
    n = len(order)
    data, gl = hbseed.hdb_pack(s_res, s_offs, n, 512, 192)
    spec = hbseed.blastp_query(hbseed.random_protein(2101, 500))
    a = hbseed.Database(ctx, s_res, s_offs)
    b = hbseed.Database.from_hdb(ctx, data, 512, 192, n, gl)
    try:
        sa, ha = hbseed.search(ctx, a, spec)
        sb, hb = hbseed.search(ctx, b, spec)
    finally:
        a.close()
        b.close()
    assert len(ha) > 100
    assert np.array_equal(ha.view(np.int32), hb.view(np.int32))
    assert np.array_equal(sa.view(np.uint8), sb.view(np.uint8))


def test_open_hdb_files(ctx, tmp_path):
    res, offs = sorted_db(9, 6000, 1, 500)
    n0 = 4000
    prefix = str(tmp_path / "vol")
    hbseed.write_hdb(prefix, res, offs, volume_seqs=[n0, 6000 - n0], blocksx=8, threadsx=64)
    dbs = hbseed.open_hdb(ctx, prefix + ".hDbI")
    try:
        assert [d.oid_base for d in dbs] == [0, n0]
        first = 0
        for d, vs in zip(dbs, [n0, 6000 - n0]):
            gs = hbseed.hdb_gpu_seqs(vs)
            assert d.n == gs
            r, o = d.fetch()
            lo = int(offs[first])
            want = offs[first:first + gs + 1] - offs[first]
            assert np.array_equal(o, want)
            assert np.array_equal(r, res[lo:lo + int(want[-1])])
            first += vs
    finally:
        for d in dbs:
            d.close()


def test_loader_rejects_inconsistent_volumes(ctx):
    res, offs = sorted_db(3, 500, 1, 100)
    data, gl = hbseed.hdb_pack(res, offs, 500, 2, 64)
    cases = [
        (data[:-16], 2, 64, 500, gl),            # short file
        (data, 2, 64, 500, gl + 1),              # group length not a multiple of 16
        (data, 2, 64, 600, gl),                  # too many sequences for the groups
        (data, 2, 32, 500, gl),                  # other grid
    ]
    for d, bx, tx, n, g in cases:
        with pytest.raises(hbseed.HBError) as e:
            hbseed.Database.from_hdb(ctx, d, bx, tx, n, g)
        assert e.value.code == -9
    bad = data.copy()
    bad[0] = 200   # residue code >= 28
    with pytest.raises(hbseed.HBError) as e:
        hbseed.Database.from_hdb(ctx, bad, 2, 64, 500, gl)
    assert e.value.code == -5


def test_loader_full_c1_volume_round_trip(ctx):
    # 2,000,000 subjects (the 1000-residue-query database), length-sorted,
    # default grid 512 x 192: the unpacked database equals the sorted arrays
    res, offs = hbseed.make_db(1100, 2_000_000)
    s_res, s_offs, order = hbseed.sort_by_length(res, offs)
    del res
    n = len(order)
    data, gl = hbseed.hdb_pack(s_res, s_offs, n, 512, 192)
    print(f"c1 volume: {n} sequences, {data.size} bytes, {gl.size} groups", flush=True)
    db = hbseed.Database.from_hdb(ctx, data, 512, 192, n, gl)
    try:
        r, o = db.fetch()
        lens_ms, copy_ms = db.hdb_kernel_ms()
    finally:
        db.close()
    print(f"kernels: lengths {lens_ms:.3f} ms, copy {copy_ms:.3f} ms", flush=True)
    sys.stdout.flush()
    assert np.array_equal(o, s_offs)
    assert np.array_equal(r, s_res)


def test_pinned_sources_load_the_same_database(ctx):
    res, offs = sorted_db(21, 4000, 1, 700)
    n = len(offs) - 1
    data, gl = hbseed.hdb_pack(res, offs, n, 4, 192)
    with ctx.pinned(data.size) as pv, ctx.pinned(res.size) as pr:
        pv.array[:] = data
        pr.array[:] = res
        a = hbseed.Database.from_hdb(ctx, pv.array, 4, 192, n, gl)
        b = hbseed.Database(ctx, pr.array, offs)
        try:
            ra, oa = a.fetch()
            rb, ob = b.fetch()
        finally:
            a.close()
            b.close()
    assert np.array_equal(oa, offs) and np.array_equal(ra, res)
    assert np.array_equal(ob, offs) and np.array_equal(rb, res)
