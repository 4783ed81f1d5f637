"""Host logic of the co-scheduled volume loop (hbseed.cosearch), no GPU: the
device-side classes are replaced by fakes that unpack a share with the
oracle's .hDb reader and scan it with the oracle, and report a device time
the test chooses.  Checks the order of the loop (volume v + 1's share is
launched before volume v's CPU search), the shares and ratios against a
replay through the oracle's balancer, and the device-budget route.
"""
import numpy as np
import pytest

import oracle_lib as O

hbseed = pytest.importorskip("hbseed")

SIZES = [3000, 2000, 2500, 2500]
BX, TX = 4, 64


class FakeCtx:
    def __init__(self, free):
        self.free = free

    def mem_info(self):
        return self.free, self.free


def fakes(monkeypatch, events, gpu_ms):
    class DB:
        def __init__(self, r, o, n):
            self.r, self.o, self.n = r, o, n

        @classmethod
        def from_hdb(cls, ctx, data, bx, tx, n, gl):
            r, o = O.hdb_unpack(np.frombuffer(data, np.uint8), n, bx, tx, np.asarray(gl))
            return cls(r, o, n)

        def close(self):
            pass

    class Query:
        def __init__(self, ctx, spec, **kw):
            self.oq = O.OracleQuery.from_spec(spec)

        def close(self):
            pass

    class Scan:
        count = 0

        def __init__(self, ctx, db, q):
            self.db, self.q = db, q
            self.k = Scan.count
            Scan.count += 1
            events.append(("launch", self.k))

        def results(self):
            cnt, h = self.q.oq.scan(self.db.r, self.db.o, threads=2)
            st = np.zeros(self.db.n, hbseed.STATS_DTYPE)
            st["lookup_hits"], st["extensions"], st["n_hsp"] = cnt[:, 0], cnt[:, 1], cnt[:, 2]
            return st, h

        def elapsed_ms(self):
            return gpu_ms[self.k]

        def close(self):
            pass

    monkeypatch.setattr(hbseed, "Database", DB)
    monkeypatch.setattr(hbseed, "Query", Query)
    monkeypatch.setattr(hbseed, "Scan", Scan)


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


def run(v, monkeypatch, gpu_ms, cpu_delay, free=1 << 40, **kw):
    import time
    events, parts = [], {}
    fakes(monkeypatch, events, gpu_ms)
    res, offs = v["res"], v["offs"]
    oq = O.OracleQuery.from_spec(v["spec"])

    def cpu_search(k, lo, hi):
        events.append(("cpu", k))
        _, h = oq.scan(res[int(offs[lo]):int(offs[hi])], offs[lo:hi + 1] - offs[lo], threads=2)
        h["oid"] += np.uint32(lo)
        parts[k] = (lo, hi, h)
        time.sleep(cpu_delay[k])

    out = hbseed.cosearch(FakeCtx(free), v["index"], v["spec"], cpu_search, 4, **kw)
    got = np.concatenate(sum(([r["hsps"], parts[k][2]] for k, r in enumerate(out)), []))
    assert np.array_equal(got.view(np.int32), v["ref"].view(np.int32))
    return out, events


def test_loop_order_and_balancer_replay(volumes, monkeypatch):
    # device 1 ms per share; the CPU 40 ms on volume 0 and 1 ms later: the
    # ratio rises after the slow CPU volume and falls back after fast ones
    out, events = run(volumes, monkeypatch, [1.0] * 4, [0.04, 0.001, 0.001, 0.001])
    launches = [k for e, k in events if e == "launch"]
    assert launches == [0, 1, 2, 3]
    # share v + 1 is launched before the CPU search of volume v
    assert events.index(("launch", 1)) < events.index(("cpu", 0))
    assert events.index(("launch", 3)) < events.index(("cpu", 2))
    o = O.runtime_options()
    ratio, fixed = o["gpu_ratio"], o["fixed_ratio"]
    rb, rr, rh = [0, 0, 0], [o["gpu_ratio"], 0.0], [0, 0]
    vols = volumes["idx"]["volumes"]
    n = int(np.asarray(volumes["spec"].seq).size)
    for v in range(-1, len(vols)):
        if v >= 0:
            rh[v & 1] = int(out[v]["hsps"].size)
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


def test_device_budget_routes_whole_volumes_to_the_cpu(volumes, monkeypatch):
    vols = volumes["idx"]["volumes"]
    opts = dict(hbseed.runtime_options(), gpu_ratio=1.0, fixed_ratio=1)
    need = [2 * v["bytes"] + 64 * v["gpu_seqs"] for v in vols]
    limit = sorted(need)[1]          # the two smallest shares fit
    out, events = run(volumes, monkeypatch, [1.0] * 4, [0.0] * 4, options=opts,
                      device_budget=limit)
    for rec, nd, vol in zip(out, need, vols):
        assert rec["on_cpu"] == (nd > limit)
        assert rec["gpu_seqs"] == (0 if nd > limit else vol["gpu_seqs"])
    assert len([e for e in events if e[0] == "launch"]) == 2
    # the default budget is the device's free memory
    out, _ = run(volumes, monkeypatch, [1.0] * 4, [0.0] * 4, free=0, options=opts)
    assert all(r["on_cpu"] for r in out)
