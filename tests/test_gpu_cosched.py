"""GPU parity of the co-scheduled volume loop (hb_cosearch, the order of
GPU_BLAST_PreliminarySearchEngine, ncbi_blast_files/functions.cpp:1409-1520):
each volume's GPU share, a prefix of its .hDb cut by hb_hdb_split, is scanned
on the device while the rest of the volume is searched on the host -- here
by the oracle, standing in for NCBI's CPU threads.  The union of both sides
equals the oracle over the whole database, HSP for HSP and counter for
counter, and the shares and ratios replay through the oracle's balancer from
the times the run recorded.
"""
import numpy as np
import pytest

import hbseed
import oracle_lib as O

pytestmark = pytest.mark.gpu

SIZES = [12000, 8000, 10000, 10000]
BLOCKSX, THREADSX = 8, 64


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    pq = hbseed.random_protein(4101, 400)
    res, offs = hbseed.make_db(4100, sum(SIZES), plant_query=pq, plant_frac=0.05)
    # volumes of consecutive oids, each sorted by length as makeblastdb's
    # volumes are for H-BLAST
    parts_r, parts_o, first = [], [np.zeros(1, np.uint64)], 0
    for vs in SIZES:
        o = offs[first:first + vs + 1] - offs[first]
        r = res[int(offs[first]):int(offs[first + vs])]
        sr, so, _ = hbseed.sort_by_length(r, o)
        parts_r.append(sr)
        parts_o.append(so[1:] + parts_o[-1][-1])
        first += vs
    res, offs = np.concatenate(parts_r), np.concatenate(parts_o)
    prefix = str(tmp_path_factory.mktemp("cosched") / "db")
    idx = hbseed.write_hdb(prefix, res, offs, volume_seqs=SIZES, blocksx=BLOCKSX,
                           threadsx=THREADSX)
    spec = hbseed.blastp_query(pq)
    oq = O.OracleQuery.from_spec(spec)
    cnt, ref = oq.scan(res, offs, threads=8)
    ctx = hbseed.Context(0)
    yield dict(res=res, offs=offs, index=prefix + ".hDbI", idx=idx, spec=spec, oq=oq, cnt=cnt,
               ref=ref, ctx=ctx)
    ctx.close()


def run(s, cpu_threads, options=None, device_budget=None):
    res, offs, oq = s["res"], s["offs"], s["oq"]
    cpu_parts = {}

    def cpu_search(v, lo, hi):
        o = offs[lo:hi + 1] - offs[lo]
        _, h = oq.scan(res[int(offs[lo]):int(offs[hi])], o, threads=4)
        h["oid"] += np.uint32(lo)
        cpu_parts[v] = (lo, hi, h)

    print(f"cosearch {len(SIZES)} volumes, {cpu_threads} CPU threads, options {options}",
          flush=True)
    out = hbseed.cosearch(s["ctx"], s["index"], s["spec"], cpu_search, cpu_threads,
                          options=options, device_budget=device_budget)
    return out, cpu_parts


def check_union(s, out, cpu_parts):
    pieces = []
    for v, rec in enumerate(out):
        base, gs = rec["oid_base"], rec["gpu_seqs"]
        lo, hi, h = cpu_parts[v]
        assert (lo, hi) == (base + gs, base + SIZES[v])
        st = rec["stats"]
        assert st.size == gs
        want = s["cnt"][base:base + gs]
        assert np.array_equal(st["lookup_hits"], want[:, 0])
        assert np.array_equal(st["extensions"], want[:, 1])
        assert np.array_equal(st["n_hsp"], want[:, 2])
        pieces += [rec["hsps"], h]
    got = np.concatenate(pieces)
    assert got.size == s["ref"].size and got.size > 500
    assert np.array_equal(got.view(np.int32), s["ref"].view(np.int32))


def test_adaptive_split_union_equals_whole_database(setup):
    s = setup
    out, cpu_parts = run(s, 4)
    for rec in out:
        print(f"  oid {rec['oid_base']}: GPU {rec['gpu_seqs']} seqs at ratio {rec['ratio']:.4f}, "
              f"gpu {rec['gpu_seconds'] * 1e3:.2f} ms, cpu {rec['cpu_seconds'] * 1e3:.2f} ms",
              flush=True)
    check_union(s, out, cpu_parts)

    # replay the balancer through the oracle with the recorded times
    vols = s["idx"]["volumes"]
    o = O.runtime_options()
    ratio, fixed = o["gpu_ratio"], o["fixed_ratio"]
    rb, rr, rh = [0, 0, 0], [o["gpu_ratio"], 0.0], [0, 0]
    seq = np.asarray(s["spec"].seq)
    for v in range(-1, len(vols)):
        if v >= 0:
            rh[v & 1] = int(out[v]["hsps"].size)
        if v + 1 < len(vols):
            nv, vol = v + 1, vols[v + 1]
            rb[nv % 3] = int(vol["bytes"]) + int(vol["cpu_rest_bytes"])
            if v > 0 and fixed != 1:
                ratio = O.next_ratio(o, v, ratio, 4, rb, rr, rh, out[v]["gpu_seconds"],
                                     out[v - 1]["cpu_seconds"], False, int(seq.size),
                                     int(seq.size))
            share, ratio, fixed = O.split(vol["group_len"], vol["gpu_seqs"], vol["bytes"],
                                          BLOCKSX, THREADSX, 4, ratio, fixed)
            rr[nv & 1] = ratio
            assert out[nv]["gpu_seqs"] == share["gpu_seqs"], nv
            assert out[nv]["ratio"] == ratio, nv
    # 4 CPU threads start at 0.8 and adapt
    assert fixed == 0
    assert 0 < out[0]["gpu_seqs"] < vols[0]["gpu_seqs"]


def test_fixed_ratio_from_options_file(setup):
    s = setup
    text = ("#ratio\n0.6\n#fixed\n1\n#capability\n12\n#hits\n20\n#ecc\n0\n#balancer\n1\n"
            "#factor\n4\n#chunk\n200\n#base\n100\n")
    opts = hbseed.runtime_options(text)
    out, cpu_parts = run(s, 16, options=opts)
    check_union(s, out, cpu_parts)
    for rec, vol in zip(out, s["idx"]["volumes"]):
        share, r, f = O.split(vol["group_len"], vol["gpu_seqs"], vol["bytes"], BLOCKSX, THREADSX,
                              16, float(np.float32(0.6)), 1)
        assert rec["gpu_seqs"] == share["gpu_seqs"] and rec["ratio"] == r
        # 60% of the workload: the long groups go to the CPU, so the GPU
        # keeps well over 60% of the sequences
        assert 0.6 * vol["gpu_seqs"] < rec["gpu_seqs"] < vol["gpu_seqs"]


def test_all_on_the_gpu_and_all_on_the_cpu(setup):
    s = setup
    base = hbseed.runtime_options()
    for ratio in (1.0, 0.0):
        opts = dict(base, gpu_ratio=ratio, fixed_ratio=1)
        out, cpu_parts = run(s, 16, options=opts)
        check_union(s, out, cpu_parts)
        for rec, vol in zip(out, s["idx"]["volumes"]):
            assert rec["gpu_seqs"] == (vol["gpu_seqs"] if ratio else 0)


def test_volume_over_the_device_budget_goes_to_the_cpu(setup):
    # the reference searches a whole volume on the CPU when its GPU share
    # does not fit the device (functions.cpp:1518-1530, 1584-1585); forced
    # here from the host with a budget below the largest share's need
    s = setup
    vols = s["idx"]["volumes"]
    opts = dict(hbseed.runtime_options(), gpu_ratio=1.0, fixed_ratio=1)
    # each whole share's need as the library computes it (hb_check_memory)
    q = hbseed.Query(s["ctx"], s["spec"])
    try:
        need = [hbseed.check_memory(s["ctx"], q, v["bytes"], v["gpu_seqs"],
                                    int(v["group_len"][-1]))[0] for v in vols]
        # the estimate bounds what a load and scan of the share hold: file,
        # residues twice, 24 B of metadata and 12 B of counters per sequence
        for v, nd in zip(vols, need):
            assert nd >= 2 * v["bytes"] + 36 * v["gpu_seqs"]
    finally:
        q.close()
    big = int(np.argmax(need))
    assert sorted(need)[-2] < need[big]
    out, cpu_parts = run(s, 16, options=opts, device_budget=need[big] - 1)
    check_union(s, out, cpu_parts)
    for k, (rec, vol) in enumerate(zip(out, vols)):
        assert rec["need_bytes"] == need[k]
        assert rec["on_cpu"] == (k == big)
        assert rec["gpu_seqs"] == (0 if k == big else vol["gpu_seqs"])
    # and with no room at all, every volume is searched on the CPU
    out, cpu_parts = run(s, 16, options=opts, device_budget=0)
    check_union(s, out, cpu_parts)
    assert all(r["on_cpu"] and r["gpu_seqs"] == 0 for r in out)
