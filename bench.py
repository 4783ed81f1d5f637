#!/usr/bin/env python3
"""Benchmark of the MI355X seed-extension stage.

Default workload (one MI355X): plain single-query BLASTP, a 1,000-aa query
(residues i.i.d. from background frequencies, seed 1101) against 2,000,000
synthetic subjects (log-normal lengths, median 300, sigma 0.6, clipped
30-5,000, 5% planted homologs; seed 1100; ~718M residues, 1 byte/residue,
resident in HBM).  Word 3, T 11, window 40, X-drop 16, cutoff 41.

One step = one pass of the stage over the whole database for the query
(hb_scan_start + hb_scan_wait: lookup, two-hit tracking, extensions, HSP
records and per-subject counters in HBM).  Prints one JSON line.
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

HBM_PEAK = 8.0e12   # bytes/s, MI355X


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def workload_c1(n_subjects):
    q = hbseed.random_protein(1101, 1000)
    res, offs = hbseed.make_db(1100, n_subjects, plant_query=q, plant_frac=0.05)
    return q, res, offs


def algorithmic_bytes(n_res, n_subj, n_hsp, n_entries):
    """Bytes the stage must move per pass: residues (1 B each), subject
    metadata (8 B offset + 4 B order index), per-subject counters written
    (12 B), HSP records (32 B each), and the lookup table read once
    (256 KB cells + 4 KB presence vector + 4 B per entry)."""
    return n_res + 12 * n_subj + 12 * n_subj + 32 * n_hsp + 262144 + 4096 + 4 * n_entries


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--subjects", type=int, default=2_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run the known-size copy kernel (rocprof calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only; --gpus must be 1 "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

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

    tr0 = time.perf_counter()
    stats, hsps = last.results()
    results_ms = (time.perf_counter() - tr0) * 1e3
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
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch")

    cpu = None if args.no_cpu else cpu_baseline(q, res, offs)

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
        "hsps": int(len(hsps)),
        "lookup_hits": int(stats["lookup_hits"].sum()),
        "extensions": int(stats["extensions"].sum()),
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": traffic},
        "cpu_baseline": cpu,
    }
    if copy:
        out["copy_calibration"] = copy
    print(json.dumps(out), flush=True)
    query.close()
    db.close()
    ctx.close()


if __name__ == "__main__":
    main()
