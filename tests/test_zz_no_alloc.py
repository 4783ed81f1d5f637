"""GPU test of hb_scan_noalloc, the scan that allocates nothing: its HSPs
(sorted into hb_scan_results' (oid, s_off, q_off) order) and per-subject
counters equal those of hb_scan_start + hb_scan_results, on two calls in a
row (the second with a smaller query and a smaller database, so stale
scratch contents would show), and for a long query whose diagonal state
goes to the scratch's global-memory part.  Device buffers come from the HIP
runtime the library itself uses.
"""
import ctypes as C

import numpy as np
import pytest

import hbseed

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the HIP runtime libhbseed links."""

    def __init__(self):
        hbseed._lib()   # loads libamdhip64 first
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.L.hipDeviceSynchronize.argtypes = []

    def malloc(self, n):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(n), 4)) == 0
        return p.value

    def free(self, p):
        self.L.hipFree(p)

    def to_host(self, dst, src, n):
        assert self.L.hipMemcpy(dst.ctypes.data, src, int(n), 2) == 0   # device to host

    def fill(self, p, byte, n):
        assert self.L.hipMemset(p, byte, int(n)) == 0

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0


@pytest.fixture(scope="module")
def ctx():
    c = hbseed.Context(0)
    yield c
    c.close()


def reference(ctx, db, q):
    s = hbseed.Scan(ctx, db, q)
    try:
        stats, hsps = s.results()
        mode = s.state_mode
    finally:
        s.close()
    return stats, hsps, mode


def noalloc(ctx, hip, db, q, n_subjects, cap):
    d_h = hip.malloc(cap * 32)
    d_c = hip.malloc(3 * n_subjects * 4)
    d_n = hip.malloc(4)
    try:
        # garbage in the outputs: the entry point must write every counter
        hip.fill(d_c, 0xA5, 3 * n_subjects * 4)
        rc = hbseed._lib().hb_scan_noalloc(ctx.h, db.h, q.h, C.c_void_p(d_h), cap,
                                           C.c_void_p(d_c), C.c_void_p(d_n), None)
        assert rc == 0, hbseed._lib().hb_strerror(rc)
        hip.sync()
        n = np.zeros(1, np.uint32)
        hip.to_host(n, d_n, 4)
        n = int(n[0])
        assert n <= cap, (n, cap)
        raw = np.zeros(max(n, 1) * 8, np.int32)
        hip.to_host(raw, d_h, n * 32)
        cnt = np.zeros(3 * n_subjects, np.uint32)
        hip.to_host(cnt, d_c, 3 * n_subjects * 4)
    finally:
        hip.free(d_h)
        hip.free(d_c)
        hip.free(d_n)
    recs = raw[:n * 8].reshape(-1, 8)
    # hb_hsp = {oid, context, q_start, s_start, q_off, s_off, len, score}
    order = np.lexsort((recs[:, 4], recs[:, 5], recs[:, 0].view(np.uint32)))
    return recs[order], cnt.reshape(-1, 3)


def check(stats, hsps, recs, cnt):
    exp = hsps.view(np.int32).reshape(-1, 8)
    assert len(exp) > 0
    assert recs.shape == exp.shape, (recs.shape, exp.shape)
    assert np.array_equal(recs, exp)
    assert np.array_equal(cnt[:, 0], stats["lookup_hits"])
    assert np.array_equal(cnt[:, 1], stats["extensions"])
    assert np.array_equal(cnt[:, 2], stats["n_hsp"])


def test_noalloc_equals_scan_twice(ctx):
    hip = Hip()
    q1 = hbseed.random_protein(5101, 1000)
    res, offs = hbseed.make_db(5100, 60000, plant_query=q1, plant_frac=0.05)
    q2 = hbseed.random_protein(5102, 300)
    n2 = 8000
    res2, offs2 = res[:int(offs[n2])].copy(), offs[:n2 + 1].copy()
    res2[int(offs2[100]):int(offs2[100]) + 300] = q2   # a planted homolog of the second query
    cases = [(hbseed.blastp_query(q1), res, offs), (hbseed.blastp_query(q2), res2, offs2)]
    for spec, r, o in cases:
        print(f"noalloc: query {len(spec.seq)} residues, {len(o) - 1} subjects", flush=True)
        db = hbseed.Database(ctx, r, o)
        q = hbseed.Query(ctx, spec)
        try:
            stats, hsps, _ = reference(ctx, db, q)
            recs, cnt = noalloc(ctx, hip, db, q, len(o) - 1, max(len(hsps) + 1000, 4096))
            check(stats, hsps, recs, cnt)
        finally:
            q.close()
            db.close()


def test_noalloc_long_query_global_state(ctx):
    hip = Hip()
    spec = hbseed.blastp_query(hbseed.random_protein(5201, 40000))
    res, offs = hbseed.make_db(5200, 1500, plant_query=spec.seq[10000:13000], plant_frac=0.3)
    db = hbseed.Database(ctx, res, offs)
    q = hbseed.Query(ctx, spec)
    try:
        stats, hsps, mode = reference(ctx, db, q)
        assert mode == hbseed.STATE_GLOBAL
        print(f"noalloc: long query, {len(hsps)} HSPs", flush=True)
        recs, cnt = noalloc(ctx, hip, db, q, len(offs) - 1, len(hsps) + 1000)
        check(stats, hsps, recs, cnt)
    finally:
        q.close()
        db.close()


def test_noalloc_counts_past_capacity(ctx):
    hip = Hip()
    qs = hbseed.random_protein(5301, 500)
    res, offs = hbseed.make_db(5300, 4000, plant_query=qs, plant_frac=0.3)
    db = hbseed.Database(ctx, res, offs)
    q = hbseed.Query(ctx, hbseed.blastp_query(qs))
    try:
        stats, hsps, _ = reference(ctx, db, q)
        d_h, d_n = hip.malloc(16 * 32), hip.malloc(4)
        try:
            rc = hbseed._lib().hb_scan_noalloc(ctx.h, db.h, q.h, C.c_void_p(d_h), 16, None,
                                               C.c_void_p(d_n), None)
            assert rc == 0
            hip.sync()
            n = np.zeros(1, np.uint32)
            hip.to_host(n, d_n, 4)
            assert int(n[0]) == len(hsps) > 16
        finally:
            hip.free(d_h)
            hip.free(d_n)
    finally:
        q.close()
        db.close()
