"""Host logic of the co-scheduled volume loop (hb_cosched_run, the C++ loop
behind hb_cosearch), no GPU: the device steps are fakes that unpack a share
with the oracle's .hDb reader, scan it with the oracle, report a device
time the test chooses and a stand-in memory need.  Checks the order of the
loop (volume v + 1's share is launched before volume v's CPU search, from
the callbacks and from the reports' sequence numbers), the shares and
ratios against a replay through the oracle's balancer, the device-memory
route and that a callback's error stops the loop.
"""
import numpy as np
import pytest

import oracle_lib as O

hbseed = pytest.importorskip("hbseed")

SIZES = [3000, 2000, 2500, 2500]
BX, TX = 4, 64


class FakeDevice:
    def __init__(self, spec, gpu_ms, events, budget=1 << 60, fail_launch=None):
        self.oq = O.OracleQuery.from_spec(spec)
        self.gpu_ms, self.events, self.budget = gpu_ms, events, budget
        self.fail_launch = fail_launch
        self.share = None
        self.results = {}

    @staticmethod
    def need(share):
        return 2 * share["bytes"] + 64 * share["gpu_seqs"]

    def fits(self, v, share, max_len):
        n = self.need(share)
        return n, n <= self.budget

    def launch(self, v, data, share, bx, tx, gl):
        assert (bx, tx) == (BX, TX) and len(data) == share["bytes"]
        if self.fail_launch == v:
            return -2   # HB_ERR_NOMEM: the volume goes to the CPU
        self.events.append(("launch", v))
        r, o = O.hdb_unpack(np.frombuffer(data, np.uint8), share["gpu_seqs"], bx, tx,
                            np.asarray(gl))
        self.share = (v, r, o, share["gpu_seqs"])
        return 0

    def fetch(self, v, oid_base):
        sv, r, o, n = self.share
        assert sv == v
        self.events.append(("fetch", v))
        cnt, h = self.oq.scan(r, o, threads=2)
        h["oid"] += np.uint32(oid_base)
        self.results[v] = (cnt, h)
        self.share = None
        return self.gpu_ms[v] / 1000.0, len(h)

    def abandon(self):
        self.events.append(("abandon", -1))
        self.share = None


@pytest.fixture(scope="module")
def volumes(tmp_path_factory):
    pq = hbseed.random_protein(6101, 300)
    res, offs = hbseed.make_db(6100, sum(SIZES), plant_query=pq, plant_frac=0.05)
    parts_r, parts_o, first = [], [np.zeros(1, np.uint64)], 0
    for vs in SIZES:
        o = offs[first:first + vs + 1] - offs[first]
        sr, so, _ = hbseed.sort_by_length(res[int(offs[first]):int(offs[first + vs])], o)
        parts_r.append(sr)
        parts_o.append(so[1:] + parts_o[-1][-1])
        first += vs
    res, offs = np.concatenate(parts_r), np.concatenate(parts_o)
    prefix = str(tmp_path_factory.mktemp("cosched_host") / "db")
    idx = hbseed.write_hdb(prefix, res, offs, volume_seqs=SIZES, blocksx=BX, threadsx=TX)
    spec = hbseed.blastp_query(pq)
    _, ref = O.OracleQuery.from_spec(spec).scan(res, offs, threads=4)
    return dict(res=res, offs=offs, index=prefix + ".hDbI", idx=idx, spec=spec, ref=ref)


def run(v, gpu_ms, cpu_delay, budget=1 << 60, fail_launch=None, **kw):
    import time
    events, parts = [], {}
    dev = FakeDevice(v["spec"], gpu_ms, events, budget, fail_launch)
    res, offs = v["res"], v["offs"]
    oq = O.OracleQuery.from_spec(v["spec"])

    def cpu_search(k, lo, hi):
        events.append(("cpu", k))
        _, h = oq.scan(res[int(offs[lo]):int(offs[hi])], offs[lo:hi + 1] - offs[lo], threads=2)
        h["oid"] += np.uint32(lo)
        parts[k] = (lo, hi, h)
        time.sleep(cpu_delay[k])

    n = int(np.asarray(v["spec"].seq).size)
    out = hbseed.cosched_run(v["index"], dev, cpu_search, 4, query_length=n,
                             avg_query_length=n, **kw)
    pieces = []
    for k, r in enumerate(out):
        if r["gpu_seqs"]:
            pieces.append(dev.results[k][1])
        lo, hi, h = parts[k]
        assert (lo, hi) == (r["oid_base"] + r["gpu_seqs"], r["oid_base"] + r["volume_seqs"])
        pieces.append(h)
    got = np.concatenate(pieces)
    assert np.array_equal(got.view(np.int32), v["ref"].view(np.int32))
    return out, events


def test_loop_order_and_balancer_replay(volumes):
    # device 1 ms per share; the CPU 40 ms on volume 0 and 1 ms later: the
    # ratio rises after the slow CPU volume and falls back after fast ones
    out, events = run(volumes, [1.0] * 4, [0.04, 0.001, 0.001, 0.001])
    assert [k for e, k in events if e == "launch"] == [0, 1, 2, 3]
    # share v + 1 is launched before the CPU search of volume v, and volume
    # v's results are fetched before it
    for v in range(3):
        assert events.index(("launch", v + 1)) < events.index(("cpu", v))
        assert events.index(("fetch", v)) < events.index(("launch", v + 1))
        assert out[v + 1]["launch_seq"] < out[v]["cpu_seq"]
    o = O.runtime_options()
    ratio, fixed = o["gpu_ratio"], o["fixed_ratio"]
    rb, rr, rh = [0, 0, 0], [o["gpu_ratio"], 0.0], [0, 0]
    vols = volumes["idx"]["volumes"]
    n = int(np.asarray(volumes["spec"].seq).size)
    for v in range(-1, len(vols)):
        if v >= 0:
            rh[v & 1] = int(out[v]["n_hsp"])
        if v + 1 < len(vols):
            nv, vol = v + 1, vols[v + 1]
            rb[nv % 3] = int(vol["bytes"]) + int(vol["cpu_rest_bytes"])
            if v > 0 and fixed != 1:
                ratio = O.next_ratio(o, v, ratio, 4, rb, rr, rh, out[v]["gpu_seconds"],
                                     out[v - 1]["cpu_seconds"], False, n, n)
            share, ratio, fixed = O.split(vol["group_len"], vol["gpu_seqs"], vol["bytes"], BX, TX,
                                          4, ratio, fixed)
            rr[nv & 1] = ratio
            assert (out[nv]["gpu_seqs"], out[nv]["ratio"]) == (share["gpu_seqs"], ratio), nv
    assert out[0]["ratio"] == pytest.approx(0.8, abs=0.01)   # 4 threads start at 0.8
    assert out[2]["ratio"] > out[1]["ratio"]                 # the slow CPU volume moved it up


def test_device_budget_routes_whole_volumes_to_the_cpu(volumes):
    vols = volumes["idx"]["volumes"]
    opts = dict(hbseed.runtime_options(), gpu_ratio=1.0, fixed_ratio=1)
    need = [FakeDevice.need(dict(bytes=v["bytes"], gpu_seqs=v["gpu_seqs"])) for v in vols]
    limit = sorted(need)[1]          # the two smallest shares fit
    out, events = run(volumes, [1.0] * 4, [0.0] * 4, budget=limit, options=opts)
    for rec, nd, vol in zip(out, need, vols):
        assert rec["on_cpu"] == (nd > limit)
        assert rec["gpu_seqs"] == (0 if nd > limit else vol["gpu_seqs"])
        assert rec["need_bytes"] == nd
    assert len([e for e in events if e[0] == "launch"]) == 2
    out, _ = run(volumes, [1.0] * 4, [0.0] * 4, budget=0, options=opts)
    assert all(r["on_cpu"] for r in out)
    # an upload that runs out of memory sends its volume to the CPU too
    out, events = run(volumes, [1.0] * 4, [0.0] * 4, options=opts, fail_launch=2)
    assert out[2]["on_cpu"] and out[2]["gpu_seqs"] == 0
    assert [k for e, k in events if e == "launch"] == [0, 1, 3]


def test_a_failing_callback_stops_the_loop(volumes):
    dev = FakeDevice(volumes["spec"], [1.0] * 4, [])

    def cpu_search(k, lo, hi):
        if k == 1:
            raise KeyError("stop here")

    with pytest.raises(KeyError):
        hbseed.cosched_run(volumes["index"], dev, cpu_search, 4)
    assert ("abandon", -1) in dev.events   # the share started for volume 2 was dropped


def test_a_zero_ratio_sends_the_rest_to_the_cpu(volumes):
    """Once a volume's ratio is 0 the reference's next update divides by it
    and stops the program (functions.cpp:65-69); the loop keeps the ratio at
    0 and the CPU searches every later volume whole, so the union still
    equals the oracle.  An adaptive start ratio of 0 gives that from the
    first volume; a device far slower than the CPU drives the ratio down
    towards it."""
    opts = hbseed.runtime_options()
    opts.update(gpu_ratio=0.0, fixed_ratio=0)
    out, events = run(volumes, [5.0] * len(SIZES), [0.0] * len(SIZES), options=opts)
    assert all(r["ratio"] == 0.0 and r["gpu_seqs"] == 0 for r in out)
    assert not [e for e in events if e[0] == "launch"]
    out, _ = run(volumes, [500000.0] * len(SIZES), [0.0] * len(SIZES))
    ratios = [r["ratio"] for r in out]
    assert all(b < a for a, b in zip(ratios[1:], ratios[2:])) and ratios[-1] < 0.1, ratios
