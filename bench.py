#!/usr/bin/env python3
"""Benchmark of the MI355X seed-extension stage.

Default workload (one MI355X): plain single-query BLASTP, a 1,000-aa query
(residues i.i.d. from background frequencies, seed 1101) against 2,000,000
synthetic subjects (log-normal lengths, median 300, sigma 0.6, clipped
30-5,000, 5% planted homologs; seed 1100; ~718M residues, 1 byte/residue,
resident in HBM).  Word 3, T 11, window 40, X-drop 16, cutoff 41.

One step = one pass of the stage over the whole database for the query
(hb_scan_start + hb_scan_wait: lookup, two-hit tracking, extensions, HSP
records and per-subject counters in HBM).  `value` = subject residues/s of
those passes.  The same line carries:
* dropin: prepare + scan + results per query (the H_BLAST_pre +
  H_BLAST_get_data pair), queries/s;
* blastx: the 1,500-nt six-frame BLASTX query against the same database,
  with its own roofline;
* gapped: the preliminary gapped extension of the last pass's HSPs (device
  kernel time, extensions, saved HSPs, the oracle on a bounded sample);
* hdb_load / db_upload: the database loaded from an H-BLAST .hDb volume and
  from packed arrays;
* forms.batched: all 2,048 batched queries in one autotuned batched scan
  (the knobs it chose) against the plain per-query loop;
* forms.served: an offered-load sweep of the served stream (80% BLASTP,
  20% BLASTX) over a 20,000,000-subject database resident in HBM, up to
  saturation: completed queries/s, p50 and p99 per point;
* roofline.traffic and valu_issue: profile-derived figures (rocprofv3 PMC
  runs of the same kernel, files named in the line).
--no-forms / --no-gapped / --no-hdb / --no-cpu skip parts; --config c3 / c4
/ qlen run the batched, served and long-query workloads alone.  Prints one
JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "h-blast_amd"), os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

import hbseed  # noqa: E402
from hbseed import workloads  # noqa: E402

HBM_PEAK = 8.0e12   # bytes/s, MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_c3(n_queries):
    """2,048 BLASTP queries, log-normal lengths (median 350, sigma 0.7,
    clipped 50-3,000), residues from background frequencies (seed 1300)."""
    return workloads.batched_queries(n_queries)


def workload_c1(n_subjects):
    return workloads.plain_db(n_subjects)


def algorithmic_bytes(n_res, n_subj, n_hsp, n_entries):
    """Bytes the stage must move per pass: residues (1 B each), one subject
    record read in processing order (4 B oid, 4 B length, 8 B offset),
    per-subject counters written (12 B), HSP records (32 B each), and the
    lookup table read once (256 KB cells + 4 KB presence vector + 4 B per
    entry)."""
    return n_res + 16 * n_subj + 12 * n_subj + 32 * n_hsp + 262144 + 4096 + 4 * n_entries


def cpu_baseline(q, res, offs, seconds_target=10.0):
    """The oracle (C restatement of the CPU two-hit word finder) on a
    bounded leading sample of the same database, all host threads."""
    import oracle_lib
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    threads = max(1, min(threads, 64))
    oq = oracle_lib.OracleQuery.from_spec(hbseed.blastp_query(q))
    n_total = len(offs) - 1
    n = min(n_total, 20000)
    while True:
        t0 = time.perf_counter()
        oq.scan(res, offs[:n + 1], threads=threads)
        dt = time.perf_counter() - t0
        if n >= n_total or dt * threads >= seconds_target or dt > seconds_target / 4:
            break
        n = min(n_total, int(n * max(2.0, seconds_target / max(dt * threads, 1e-3))))
    residues = int(offs[n] - offs[0])
    return {"value": residues / dt, "unit": "subject residues/s", "cores": threads,
            "kind": "port",
            "sample": f"first {n} of {n_total} subjects ({residues} residues), "
                      f"{dt:.2f} s wall on {threads} threads"}


def gapped_cutoff(query_len, db_residues, evalue=10.0, lam=0.267, K=0.041):
    """Raw gapped cutoff for E-value `evalue`: ceil(ln(K m n / E) / lambda)
    with BLOSUM62 11/1 gapped lambda 0.267 and K 0.041, over the search space
    m x n without NCBI's length adjustment."""
    return int(np.ceil(np.log(K * query_len * db_residues / evalue) / lam))


def measure_gapped(scan, q, res, offs, hsps, cutoff, reps=3, cpu=True):
    """The preliminary gapped extension (hb_scan_gapped) on one scan's HSPs:
    device kernel time (hipEvents around the gapped kernels), the whole call
    (layout on the device, kernels, copy out), and the oracle on a bounded
    leading sample of the same HSPs.  The host arrays are allocated by a
    first, untimed call and reused."""
    kernel, call = [], []
    bufs = hbseed.ResultBuffers()
    scan.gapped([cutoff], out=bufs)
    for _ in range(reps):
        t0 = time.perf_counter()
        gst, g = scan.gapped([cutoff], out=bufs)
        call.append((time.perf_counter() - t0) * 1e3)
        ms, n_wide = scan.gapped_info()
        kernel.append(ms)
    ext = int(gst["extensions"].sum())
    out = {"kernel_ms": float(np.mean(kernel)), "call_ms": float(np.mean(call)),
           "ungapped_hsps": int(len(hsps)), "extensions": ext, "hsps": int(len(g)),
           "wide_subjects": int(n_wide),
           "extensions_per_s": ext / (np.mean(kernel) * 1e-3),
           "params": {"gap_open": 11, "gap_extend": 1, "x_dropoff": 38, "cutoff": cutoff}}
    if cpu:
        import oracle_lib
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
        threads = max(1, min(threads, 64))
        oq = oracle_lib.OracleQuery.from_spec(hbseed.blastp_query(q))
        n_total = len(offs) - 1
        n = min(n_total, 100000)
        while True:
            k = int(np.searchsorted(hsps["oid"], n))
            t0 = time.perf_counter()
            cnt, _ = oracle_lib.gapped_db(oq, res, offs[:n + 1], hsps[:k], [cutoff],
                                          threads=threads)
            dt = time.perf_counter() - t0
            if n >= n_total or dt * threads >= 10.0 or dt > 2.5:
                break
            n = min(n_total, int(n * max(2.0, 10.0 / max(dt * threads, 1e-3))))
        sample_ext = int(cnt[:, 1].sum())
        out["cpu_baseline"] = {
            "value": sample_ext / dt, "unit": "gapped extensions/s", "cores": threads,
            "kind": "port",
            "sample": f"HSPs of the first {n} of {n_total} subjects ({k} ungapped HSPs, "
                      f"{sample_ext} extensions), {dt:.2f} s wall on {threads} threads"}
    return out


def measure_hdb(ctx, res, offs, reps=3, cpu=True, blocksx=512, threadsx=192):
    """The same database loaded as an H-BLAST .hDb volume (hb_db_upload_hdb):
    the sequences sorted by length, the volume's GPU share (99.9%, a multiple
    of 32) packed on the host in the reference's layout (default grid
    512 x 192), then uploaded and unpacked on the device.  Reports the two
    unpacking kernels (hipEvents), the whole call (PCIe upload included) and
    the oracle's reader on the same file."""
    s_res, s_offs, _ = hbseed.sort_by_length(res, offs)
    n = hbseed.hdb_gpu_seqs(len(s_offs) - 1)
    data, gl = hbseed.hdb_pack(s_res, s_offs, n, blocksx, threadsx)
    n_res = int(s_offs[n])
    # file read once, residues written, lengths written and read, offsets read
    abytes = data.size + n_res + 16 * n
    # the volume in page-locked memory, as a reader would put it there
    pin = ctx.pinned(data.size)
    pin.array[:] = data
    call, call_pinned, k_len, k_copy = [], [], [], []
    for it in range(reps + 1):   # the first load also loads the kernels' code: untimed
        for src, calls in ((data, call), (pin.array, call_pinned)):
            t0 = time.perf_counter()
            db = hbseed.Database.from_hdb(ctx, src, blocksx, threadsx, n, gl)
            dt = (time.perf_counter() - t0) * 1e3
            a, b = db.hdb_kernel_ms()
            if it:
                calls.append(dt)
                k_len.append(a)
                k_copy.append(b)
            ok = db.n == n and db.residues == n_res
            db.close()
            if not ok:
                raise RuntimeError("hb_db_upload_hdb: loaded volume has the wrong size")
    pin.close()
    kms = float(np.mean(k_len) + np.mean(k_copy))
    achieved = abytes / (kms * 1e-3)
    out = {"volume": f"{n} shortest of {len(s_offs) - 1} subjects, grid {blocksx}x{threadsx}, "
                     f"{gl.size} groups",
           "sequences": n, "file_bytes": int(data.size), "residues": n_res,
           "lengths_kernel_ms": float(np.mean(k_len)), "copy_kernel_ms": float(np.mean(k_copy)),
           "call_ms": float(np.mean(call)),
           "call_gb_per_s": data.size / (np.mean(call) * 1e-3) / 1e9,
           "call_pinned_ms": float(np.mean(call_pinned)),
           "call_pinned_gb_per_s": data.size / (np.mean(call_pinned) * 1e-3) / 1e9,
           "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": None}}
    if cpu:
        import oracle_lib
        t0 = time.perf_counter()
        r2, o2 = oracle_lib.hdb_unpack(data, n, blocksx, threadsx, gl)
        dt = time.perf_counter() - t0
        if int(o2[-1]) != n_res:
            raise RuntimeError("oracle reader disagrees on the volume size")
        out["cpu_baseline"] = {
            "value": abytes / dt / 1e9, "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"the whole {data.size}-byte volume, {dt:.2f} s on one thread"}
    return out


def measure_upload(ctx, res, offs, reps=2):
    """hb_db_upload of the packed database from ordinary (pageable) memory and
    from page-locked memory (hb_host_alloc): wall time of the whole call
    (host-to-device copy, processing order, residue check)."""
    pin = ctx.pinned(res.size)
    pin.array[:] = res
    out = {"bytes": int(res.size)}
    for key, src in (("pageable_ms", res), ("pinned_ms", pin.array)):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            d = hbseed.Database(ctx, src, offs)
            ts.append((time.perf_counter() - t0) * 1e3)
            d.close()
        out[key] = min(ts)
    pin.close()
    out["pinned_gb_per_s"] = res.size / (out["pinned_ms"] * 1e-3) / 1e9
    return out


def measure_batched(ctx, db, qs, steps, warmup, plain_sample, group_residues=0):
    """All queries in one batched scan (groups of group_residues, 0 =
    autotuned; hipGraph replay), beside the plain per-query loop (prepare +
    scan + results per query) on the first plain_sample queries.  One step
    = one batched pass over the database for all queries."""
    b = hbseed.Batch(ctx, db, group_residues=group_residues)
    for q in qs:
        b.add(hbseed.blastp_query(q))
    t0 = time.perf_counter()
    b.run()    # builds groups + lookup tables (autotuned), captures the graph
    first_ms = (time.perf_counter() - t0) * 1e3
    log(f"batched first run {first_ms:.0f} ms: {b.info()}")
    for _ in range(max(0, warmup - 1)):
        b.run()
    ctx.synchronize()
    kms = []
    t0 = time.perf_counter()
    for k in range(steps):
        b.run()
        kms.append(b.info()["kernel_ms"])
        log(f"batched step {k + 1}/{steps}: kernel {kms[-1]:.0f} ms")
    ctx.synchronize()
    wall = time.perf_counter() - t0
    info = b.info()
    b.close()
    sample = qs[:plain_sample]
    for q in sample[:2]:
        hbseed.search(ctx, db, hbseed.blastp_query(q))
    ctx.synchronize()
    tp = time.perf_counter()
    for q in sample:
        hbseed.search(ctx, db, hbseed.blastp_query(q))
    ctx.synchronize()
    plain_qps = len(sample) / (time.perf_counter() - tp)
    qps = len(qs) * steps / wall
    return {"queries": len(qs), "query_residues": int(sum(len(q) for q in qs)),
            "queries_per_s": qps, "plain_queries_per_s": plain_qps,
            "plain_sample": len(sample), "speedup_vs_plain": qps / plain_qps,
            "ms_per_step": wall / steps * 1e3, "kernel_ms": float(np.mean(kms)),
            "first_run_ms": first_ms, "build_ms": info["build_ms"],
            "group_residues": info["group_residues"], "groups": info["n_groups"],
            "autotuned": {"group_residues": info["group_residues"],
                          "waves_per_workgroup": info["waves_per_workgroup"],
                          "ring_bytes": info["ring_bytes"],
                          "items_per_workgroup": info["items_per_workgroup"],
                          "candidates_timed": info["tune_candidates"]} if group_residues == 0
            else None,
            "launches": info["n_classes"], "hsps": info["hsps"],
            "lookup_hits": info["lookup_hits"], "extensions": info["extensions"]}


def measure_served(ctx, db, specs, arrivals, max_batch_residues, max_inflight, check=3,
                   group_residues=None):
    """Submits specs[i] at time arrivals[i] (seconds) from a client thread to
    a server over the resident database, waits for every request, and
    returns throughput and latency (submit -> results ready, from the
    server's own clock).  `check` sampled results are compared with the
    plain per-query path."""
    import threading
    n = len(specs)
    srv = hbseed.Server(ctx, db, max_batch_residues=max_batch_residues,
                        max_inflight=max_inflight, group_residues=group_residues or None)
    # warm-up request (first launch, code-object load), not counted
    t = srv.submit(hbseed.blastp_query(hbseed.random_protein(1499, 300)))
    srv.wait(t)
    srv.results(t)
    st0 = srv.stats()
    tickets = [None] * n
    done_t = [0.0] * n
    n_hsp = [0] * n
    sample = set(np.linspace(0, n - 1, min(check, n)).astype(int).tolist())
    kept = {}
    t0 = time.perf_counter()

    def client():
        for i in range(n):
            d = t0 + arrivals[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(specs[i][1])

    th = threading.Thread(target=client)
    th.start()
    last_log = time.perf_counter()
    for i in range(n):
        while tickets[i] is None:
            time.sleep(0.001)
        srv.wait(tickets[i])
        done_t[i] = time.perf_counter()
        h = srv.results(tickets[i])
        n_hsp[i] = len(h)
        if i in sample:
            kept[i] = h
        if time.perf_counter() - last_log > 20:
            log(f"served {i + 1}/{n}")
            last_log = time.perf_counter()
    th.join()
    wall = max(done_t) - t0
    st = srv.stats()
    srv.stop()
    ok = True
    for i, h in kept.items():
        _, ref = hbseed.search(ctx, db, specs[i][1])
        ok = ok and np.array_equal(h.view(np.int32), ref.view(np.int32))
    if kept:
        log(f"served results equal plain path on {len(kept)} sampled queries: {ok}")
    n_done = st["completed"] - st0["completed"]
    return {"arrivals": n, "blastx": sum(1 for k, _ in specs if k == "blastx"),
            "completed": n_done, "queries_per_s": n_done / wall, "wall_s": wall,
            "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
            "max_ms": st["max_ms"],
            "query_residues_per_s": (st["query_residues"] - st0["query_residues"]) / wall,
            "batches": st["batches"] - st0["batches"], "peak_inflight": st["peak_inflight"],
            "reruns": st["reruns"], "gpu_ms": st["gpu_ms"] - st0["gpu_ms"],
            "hsps": int(sum(n_hsp)), "sample_equal_plain": ok}


def profile_figures():
    """Figures read from the committed rocprofv3 summaries of the scan
    kernel on the plain workload (profile-derived, with their source): HBM
    traffic per launch (FETCH_SIZE + WRITE_SIZE, scaled by a known-size copy)
    and the VALU issue fraction (SQ_INSTS_VALU x 4 cycles / (SIMDs x clock x
    kernel time))."""
    out = {}
    for key, name in (("traffic", "hb_scan_kernel_c1_traffic.json"),
                      ("issue", "hb_scan_kernel_c1_sq_counters.json")):
        path = os.path.join(ROOT, "profiles", name)
        if os.path.exists(path):
            with open(path) as f:
                out[key] = dict(json.load(f), source=f"profiles/{name}")
    return out


def measure_dropin(ctx, db, spec, reps=10):
    """The drop-in per query as H-BLAST's H_BLAST_pre + H_BLAST_get_data
    pair does it: prepare (lookup table on the device), scan, and results
    (per-subject stats of every subject and the ordered HSPs copied to the
    host), after one warm-up; mean wall time per query.  The host result
    arrays are allocated by the warm-up and reused, as H-BLAST's caller keeps
    its result buffers across volumes (functions.cpp:218-235).  The same loop
    with the compact results (entries of the subjects with HSPs, totals) is
    reported beside it."""
    bufs = {False: hbseed.ResultBuffers(), True: hbseed.ResultBuffers()}

    def one(compact):
        q = hbseed.Query(ctx, spec)
        s = hbseed.Scan(ctx, db, q)
        r = s.results_compact(bufs[True]) if compact else s.results(bufs[False])
        s.close()
        q.close()
        return len(r[1])
    out = {}
    for compact in (False, True):
        one(compact)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            one(compact)
        ctx.synchronize()
        dt = (time.perf_counter() - t0) / reps
        if compact:
            out["compact"] = {"queries_per_s": 1.0 / dt, "ms_per_query": dt * 1e3,
                              "steps": "prepare + scan + compact results (subjects with "
                                       "HSPs, totals, HSPs to the host)"}
        else:
            out.update({"queries_per_s": 1.0 / dt, "ms_per_query": dt * 1e3, "reps": reps,
                        "steps": "prepare (lookup table) + scan + results (stats of every "
                                 "subject and HSPs to the host)"})
    return out


def measure_blastx(ctx, db, res, offs, reps=5):
    """The plain BLASTX workload: the 1,500-nt six-frame query against the
    same database; device time of the scan and its roofline."""
    spec = workloads.blastx_c2(res, offs)
    q = hbseed.Query(ctx, spec)
    n_entries = q.lut()[1].size
    ms = []
    for k in range(reps + 1):
        s = hbseed.Scan(ctx, db, q)
        s.wait()
        if k:
            ms.append(s.elapsed_ms())
        if k < reps:
            s.close()
    st, h = s.results()
    s.close()
    q.close()
    n_res, n_subj = int(offs[-1]), len(offs) - 1
    kms = float(np.mean(ms))
    abytes = algorithmic_bytes(n_res, n_subj, len(h), n_entries)
    traffic = src = None
    tf = os.path.join(ROOT, "profiles", "hb_scan_kernel_c2_traffic.json")
    if os.path.exists(tf):   # profile-derived: rocprofv3 PMC runs of these launches
        with open(tf) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")
        src = "profiles/hb_scan_kernel_c2_traffic.json"
    return {"workload": "plain BLASTX: 1,500-nt query (six frames, two planted coding regions) "
                        "vs the same 2,000,000 subjects",
            "query_residues": int(len(spec.seq)), "contexts": int(spec.n_ctx),
            "kernel_ms": kms, "subject_residues_per_s": n_res / (kms * 1e-3),
            "hsps": int(len(h)), "frames_with_hsps": int(len(np.unique(h["context"]))),
            "lookup_hits": int(st["lookup_hits"].sum()),
            "extensions": int(st["extensions"].sum()),
            "roofline": {"bound": "hbm", "achieved": abytes / (kms * 1e-3) / 1e9,
                         "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                         "frac": abytes / (kms * 1e-3) / HBM_PEAK, "traffic": traffic,
                         "traffic_source": src}}


def measure_served_sweep(ctx, db, n_res, n_subj, start_qps=2.0, factor=1.6, points=6,
                         seconds=8.0):
    """Offered-load sweep of the served form: Poisson arrivals of the served
    mix at rising rates, seconds of arrivals each, until the completed rate
    falls below 85% of the offered one (saturation) or p99 passes 4 s."""
    gaps, specs = workloads.served_stream(4000)
    rows = []
    qps = start_qps
    for k in range(points):
        n_arr = min(len(specs), int(qps * seconds) + 1)
        arrivals = np.cumsum(gaps[:n_arr] / qps)
        log(f"served sweep: {qps:.2f} q/s offered, {n_arr} arrivals")
        m = measure_served(ctx, db, specs[:n_arr], arrivals, 16384, 2, check=2 if k == 0 else 0)
        row = {"offered_qps": qps, "completed_qps": m["queries_per_s"],
               "p50_ms": m["p50_ms"], "p99_ms": m["p99_ms"], "mean_ms": m["mean_ms"],
               "arrivals": m["arrivals"], "blastx": m["blastx"], "batches": m["batches"],
               "peak_inflight": m["peak_inflight"],
               # each batch's event-timed span summed over the wall time: above 1
               # when the batches in flight overlap on the device
               "batch_span_per_wall": m["gpu_ms"] / 1e3 / m["wall_s"],
               "effective_subject_residues_per_s": m["completed"] * n_res / m["wall_s"]}
        if k == 0:
            row["sample_equal_plain"] = m["sample_equal_plain"]
        rows.append(row)
        if m["queries_per_s"] < 0.85 * qps or m["p99_ms"] > 4000:
            break
        qps *= factor
    best = max(rows, key=lambda r: r["completed_qps"])
    return {"workload": f"served stream: Poisson arrivals, 80% BLASTP (log-normal median 350) / "
                        f"20% BLASTX (600-3000 nt) vs {n_subj} subjects ({n_res} residues) "
                        "resident",
            "subjects": n_subj, "residues": n_res, "points": rows,
            "max_completed_qps": best["completed_qps"],
            "p50_ms_at_max": best["p50_ms"], "p99_ms_at_max": best["p99_ms"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--subjects", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the known-size copy kernel (rocprof calibration)")
    ap.add_argument("--config", default="c1", choices=["c1", "c3", "c4", "qlen"],
                    help="c1: plain 1000-aa query (default); c3: batched 2,048 queries; "
                         "c4: served Poisson stream over 20M subjects; qlen: plain scans "
                         "of long queries with the diagonal state in LDS and in global "
                         "memory")
    ap.add_argument("--queries", type=int, default=2048, help="c3: number of queries")
    ap.add_argument("--group-residues", type=int, default=0,
                    help="c3: query residues per group (0: autotuned); c4: per served group "
                         "(0: the server's default)")
    ap.add_argument("--plain-sample", type=int, default=48,
                    help="c3: queries timed one by one through the plain path")
    ap.add_argument("--qps", type=float, default=4.0, help="c4: Poisson arrival rate")
    ap.add_argument("--duration", type=float, default=40.0,
                    help="c4: seconds of arrivals")
    ap.add_argument("--serve-subjects", type=int, default=20_000_000,
                    help="c4: database subjects")
    ap.add_argument("--max-batch-residues", type=int, default=16384,
                    help="c4: query residues per served batch")
    ap.add_argument("--max-inflight", type=int, default=2, help="c4: batches in flight")
    ap.add_argument("--no-forms", action="store_true",
                    help="c1: skip the batched (all queries) and served (sweep) measurements")
    ap.add_argument("--no-gapped", action="store_true",
                    help="c1: skip the gapped-extension measurement")
    ap.add_argument("--no-hdb", action="store_true",
                    help="c1: skip loading the database as an .hDb volume")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.config == "c3":
        return bench_c3(args)
    if args.config == "c4":
        return bench_c4(args)
    if args.config == "qlen":
        return bench_qlen(args)
    log("generating workload")
    q, res, offs = workload_c1(args.subjects)
    n_subj = len(offs) - 1
    n_res = int(offs[-1])
    log(f"{n_subj} subjects, {n_res} residues")

    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    query = hbseed.Query(ctx, hbseed.blastp_query(q))
    n_entries = query.lut()[1].size
    log("database resident; warming up")
    for _ in range(args.warmup):
        s = hbseed.Scan(ctx, db, query)
        s.wait()
        s.close()

    ctx.synchronize()
    kernel_ms = []
    t0 = time.perf_counter()
    last = None
    for k in range(args.steps):
        s = hbseed.Scan(ctx, db, query)
        s.wait()
        kernel_ms.append(s.elapsed_ms())
        if last is not None:
            last.close()
        last = s
    ctx.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    log(f"timed {args.steps} steps in {wall:.3f} s")

    # first fetch: includes the context's one-time pinned staging allocation
    # and the allocation and first touch of the host arrays; the second, into
    # the same arrays, is the per-query cost
    bufs = hbseed.ResultBuffers()
    tr0 = time.perf_counter()
    stats, hsps = last.results(bufs)
    results_first_ms = (time.perf_counter() - tr0) * 1e3
    tr0 = time.perf_counter()
    stats, hsps = last.results(bufs)
    results_ms = (time.perf_counter() - tr0) * 1e3
    stats, hsps = stats.copy(), hsps.copy()
    # the compact form: entries of the subjects with HSPs and the totals
    cbufs = hbseed.ResultBuffers()
    last.results_compact(cbufs)
    tr0 = time.perf_counter()
    subj_c, _, _ = last.results_compact(cbufs)
    results_compact_ms = (time.perf_counter() - tr0) * 1e3
    gapped = None
    if not args.no_gapped:
        log("gapped extension of the last pass's HSPs")
        gapped = measure_gapped(last, q, res, offs, hsps, gapped_cutoff(len(q), n_res),
                                cpu=not args.no_cpu)
    last.close()

    copy = None
    if args.calibrate_copy:
        import ctypes as C
        ms = C.c_float()
        nbytes = 1 << 30
        rc = hbseed._lib().hb_diag_copy(ctx.h, nbytes, 5, C.byref(ms))
        if rc == 0:
            copy = {"bytes": nbytes, "ms": ms.value,
                    "gbps": 2 * nbytes / (ms.value * 1e-3) / 1e9}

    mean_kernel_ms = float(np.mean(kernel_ms))
    abytes = algorithmic_bytes(n_res, n_subj, len(hsps), n_entries)
    achieved = abytes / (mean_kernel_ms * 1e-3)
    traffic = None
    tf = os.path.join(ROOT, "profiles", "hb_scan_kernel_c1_traffic.json")
    if os.path.exists(tf):   # profile-derived: rocprofv3 PMC runs of this kernel
        with open(tf) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    cpu = None if args.no_cpu else cpu_baseline(q, res, offs)

    hdb = upload = None
    if not args.no_hdb:
        log("database as an .hDb volume")
        hdb = measure_hdb(ctx, res, offs, cpu=not args.no_cpu)
        log("database upload from pageable and page-locked memory")
        upload = measure_upload(ctx, res, offs)

    log("drop-in per query (prepare + scan + results)")
    dropin = measure_dropin(ctx, db, hbseed.blastp_query(q))
    log("plain BLASTX")
    blastx = measure_blastx(ctx, db, res, offs)
    prof = profile_figures()

    # the batched workload (all queries) on the same resident database, and
    # the served stream on its own 20,000,000-subject database
    forms = None
    if not args.no_forms:
        forms = {}
        try:
            log(f"batched: {args.queries} queries")
            bat = measure_batched(ctx, db, workload_c3(args.queries), steps=1, warmup=1,
                                  plain_sample=args.plain_sample)
            bat["workload"] = (f"batched BLASTP: {args.queries} queries (log-normal median 350, "
                               "sigma 0.7, clipped 50-3000) vs the same 2,000,000 subjects")
            forms["batched"] = bat
        except Exception as e:   # reported in the line; the plain figure stands
            forms["batched"] = {"error": repr(e)}
        try:
            log(f"served: {args.serve_subjects} subjects")
            sres, soffs = workloads.served_db(args.serve_subjects)
            sdb = hbseed.Database(ctx, sres, soffs)
            n_sres, n_ssubj = int(soffs[-1]), len(soffs) - 1
            del sres, soffs
            forms["served"] = measure_served_sweep(ctx, sdb, n_sres, n_ssubj)
            sdb.close()
        except Exception as e:
            forms["served"] = {"error": repr(e)}

    out = {
        "metric": "subject residues/s",
        "value": n_res * args.steps / wall,
        "unit": "subject residues/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": "plain single-query BLASTP drop-in: 1000-aa query vs "
                               "2,000,000 subjects (log-normal median 300, sigma 0.6, "
                               "5% planted homologs)",
                   "query_len": 1000, "subjects": n_subj, "residues": n_res,
                   "word": 3, "threshold": 11, "window": 40, "xdrop": 16, "cutoff": 41,
                   "seeds": {"db": 1100, "query": 1101}},
        "queries_per_s": args.steps / wall,
        "kernel_ms": mean_kernel_ms,
        "results_fetch_ms": results_ms,
        "results_fetch_first_ms": results_first_ms,
        "results_compact_ms": results_compact_ms,
        "subjects_with_hsps": int(len(subj_c)),
        "hsps": int(len(hsps)),
        "lookup_hits": int(stats["lookup_hits"].sum()),
        "extensions": int(stats["extensions"].sum()),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "traffic_source": prof.get("traffic", {}).get("source")},
        "cpu_baseline": cpu,
        "dropin": dropin,
        "blastx": blastx,
    }
    if "issue" in prof:
        iss = prof["issue"]
        out["valu_issue"] = {k: iss.get(k) for k in ("valu_issue_frac", "valu_per_hit",
                                                      "salu_per_hit", "lds_per_hit",
                                                      "lds_bank_conflict_frac", "kernel_ms",
                                                      "source")}
    if copy:
        out["copy_calibration"] = copy
    if gapped:
        out["gapped"] = gapped
    if hdb:
        out["hdb_load"] = hdb
    if upload:
        out["db_upload"] = upload
    if forms:
        out["forms"] = forms
    print(json.dumps(out), flush=True)
    query.close()
    db.close()
    ctx.close()


def bench_qlen(args, n_subjects=200_000,
               lengths=(1000, 2000, 4000, 6000, 8000, 12000, 16000, 24000, 31000)):
    """Plain BLASTP scans of one query per length (i.i.d. residues, seed
    7000 + length) against 200,000 subjects of the default database
    generator (seed 1100), with the per-diagonal two-hit state in LDS
    (where it fits) and in global memory; the best of two timed scans after
    one warm-up, device time per scan."""
    res, offs = hbseed.make_db(1100, n_subjects)
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    rows = []
    for n in lengths:
        q = hbseed.Query(ctx, hbseed.blastp_query(hbseed.random_protein(7000 + n, n)))
        s = hbseed.Scan(ctx, db, q)
        row = {"query_len": n, "auto": {hbseed.STATE_LDS: "lds",
                                        hbseed.STATE_GLOBAL: "global"}[s.state_mode]}
        s.close()
        for mode, key in ((hbseed.STATE_LDS, "lds_ms"), (hbseed.STATE_GLOBAL, "global_ms")):
            try:
                q.set_state_mode(mode)
            except hbseed.HBError:
                row[key] = None
                continue
            ms = []
            for _ in range(3):
                s = hbseed.Scan(ctx, db, q)
                s.wait()
                ms.append(s.elapsed_ms())
                s.close()
            row[key] = min(ms[1:])
        log(f"query {n}: {row}")
        rows.append(row)
        q.close()
    out = {"metric": "ms per scan by query length", "unit": "ms", "n_gpus": 1,
           "higher_is_better": False, "dtype": "int32", "data": "synthetic",
           "config": {"workload": f"plain BLASTP scans of 1,000-31,000-residue queries vs "
                                  f"{n_subjects:,} subjects", "subjects": n_subjects,
                      "residues": int(offs[-1])},
           "rows": rows}
    print(json.dumps(out), flush=True)
    db.close()
    ctx.close()


def bench_c3(args):
    """Batched BLASTP: all queries in one batched scan (autotuned groups,
    hipGraph replay) against the C1 database, beside the plain per-query
    loop on a sample of the same queries."""
    log("generating workload")
    _, res, offs = workload_c1(args.subjects)
    qs = workload_c3(args.queries)
    n_res = int(offs[-1])
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    log(f"{len(offs) - 1} subjects resident; {len(qs)} queries, "
        f"{sum(len(q) for q in qs)} query residues")
    m = measure_batched(ctx, db, qs, args.steps, args.warmup, args.plain_sample,
                        group_residues=args.group_residues)
    abytes = n_res + 28 * (len(offs) - 1) + 32 * m["hsps"]
    out = {
        "metric": "queries/s (batched vs plain)",
        "value": m["queries_per_s"],
        "unit": "queries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"batched BLASTP: {len(qs)} queries (log-normal median 350, "
                               "sigma 0.7, clipped 50-3000) vs 2,000,000 subjects",
                   "queries": len(qs), "query_residues": m["query_residues"],
                   "subjects": len(offs) - 1, "residues": n_res,
                   "group_residues": m["group_residues"], "groups": m["groups"],
                   "launches": m["launches"], "seeds": {"db": 1100, "queries": 1300}},
        "effective_subject_residues_per_s": m["queries_per_s"] * n_res,
        "plain_queries_per_s": m["plain_queries_per_s"],
        "plain_sample": m["plain_sample"],
        "speedup_vs_plain": m["speedup_vs_plain"],
        "kernel_ms": m["kernel_ms"],
        "first_run_ms": m["first_run_ms"],
        "build_ms": m["build_ms"],
        "hsps": m["hsps"], "lookup_hits": m["lookup_hits"], "extensions": m["extensions"],
        "roofline": {"bound": "hbm", "achieved": abytes / (m["kernel_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": abytes / (m["kernel_ms"] * 1e-3) / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    db.close()
    ctx.close()


def workload_c4(n, seed=1400):
    """Arrival offsets (Poisson process at rate qps per second is applied by
    the caller) and the query mix: 80% BLASTP with lengths as the batched
    workload (log-normal median 350, sigma 0.7, clipped 50-3,000), 20%
    BLASTX of U[600, 3000] nt with two planted coding regions."""
    return workloads.served_stream(n, seed)


def bench_c4(args):
    """Served stream: Poisson arrivals at --qps for --duration seconds,
    submitted from a client thread to the server (hb_serve_*), which folds
    them into batched scans over the resident database.  Reports completed
    queries/s, subject residues/s (queries x database residues / wall) and
    p50/p99 latency from submit to results ready; a sample of the served
    results is checked against the plain per-query path."""
    log(f"generating {args.serve_subjects} subjects")
    res, offs = workloads.served_db(args.serve_subjects)
    n_res = int(offs[-1])
    n_arr = max(1, int(args.qps * args.duration * 1.5) + 20)
    gaps, specs = workload_c4(n_arr)
    arrivals = np.cumsum(gaps / args.qps)
    n_sub = max(1, int(np.searchsorted(arrivals, args.duration)))
    specs, arrivals = specs[:n_sub], arrivals[:n_sub]
    ctx = hbseed.Context(0)
    db = hbseed.Database(ctx, res, offs)
    del res
    log(f"{len(offs) - 1} subjects, {n_res} residues resident; {n_sub} arrivals over "
        f"{args.duration:.0f} s at {args.qps} q/s")
    m = measure_served(ctx, db, specs, arrivals, args.max_batch_residues, args.max_inflight,
                       group_residues=args.group_residues)
    abytes = m["batches"] * (n_res + 28 * (len(offs) - 1)) + 32 * m["hsps"]
    out = {
        "metric": "queries/s served (p50/p99 latency)",
        "value": m["queries_per_s"],
        "unit": "queries/s",
        "n_gpus": 1, "steps": n_sub, "warmup": 1,
        "ms_per_step": m["wall_s"] / max(m["completed"], 1) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"served stream: Poisson arrivals at {args.qps} q/s for "
                               f"{args.duration:.0f} s, 80% BLASTP (log-normal median 350) / "
                               f"20% BLASTX (600-3000 nt) vs {len(offs) - 1} subjects",
                   "subjects": len(offs) - 1, "residues": n_res, "arrivals": n_sub,
                   "blastx": m["blastx"], "max_batch_residues": args.max_batch_residues,
                   "max_inflight": args.max_inflight, "seeds": {"db": 1400, "stream": 1400}},
        "p50_ms": m["p50_ms"], "p99_ms": m["p99_ms"], "mean_ms": m["mean_ms"],
        "max_ms": m["max_ms"],
        "effective_subject_residues_per_s": m["completed"] * n_res / m["wall_s"],
        "query_residues_per_s": m["query_residues_per_s"],
        "batches": m["batches"], "peak_inflight": m["peak_inflight"], "reruns": m["reruns"],
        # each batch's event-timed span summed over the wall time (above 1
        # when the batches in flight overlap on the device)
        "batch_span_per_wall": m["gpu_ms"] / (m["wall_s"] * 1e3),
        "hsps": m["hsps"],
        "sample_equal_plain": m["sample_equal_plain"],
        "roofline": {"bound": "hbm", "achieved": abytes / (m["gpu_ms"] * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": abytes / (m["gpu_ms"] * 1e-3) / HBM_PEAK, "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)
    db.close()
    ctx.close()


if __name__ == "__main__":
    main()
