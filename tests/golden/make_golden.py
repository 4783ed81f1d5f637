#!/usr/bin/env python3
"""Writes the golden fixtures of the seed-extension stage.

Each fixture holds its inputs (query contexts, packed subjects) and the
oracle's outputs (HSP records in (oid, s_off, q_off) order and per-subject
counters: lookup hits, extensions, HSPs), as .npz without pickles, plus a
SHA-256 manifest.  The reference ships no golden vectors for this path;
these are the oracle's, and the oracle itself is pinned by
tests/test_oracle_pins.py.

Run from the repository root after building: python3 tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "h-blast_amd"), os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import hbseed  # noqa: E402
import oracle_lib  # noqa: E402


def cases():
    # BLASTP, 300-aa query vs 2,000 subjects (config-0 style, 10% planted)
    q = hbseed.random_protein(1001, 300)
    res, offs = hbseed.make_db(1000, 2000, plant_query=q, plant_frac=0.1)
    yield "blastp_q300_s2000", hbseed.blastp_query(q), res, offs
    # BLASTX, 600-nt query with two planted coding regions vs 1,000 subjects
    pq = hbseed.random_protein(1201, 400)
    res, offs = hbseed.make_db(1200, 1000, plant_query=pq, plant_frac=0.1)
    seg1 = hbseed.homolog(1203, res[int(offs[3]):int(offs[3]) + 60], 0.7, 1)
    seg2 = hbseed.homolog(1204, pq[100:160], 0.6, 0)
    nt = hbseed.blastx_nt(1202, 600, seg1, seg2)
    yield "blastx_nt600_s1000", hbseed.blastx_query(nt), res, offs
    # edge subjects: empty, 1-3 residues, query copies, a run, a long subject
    q = hbseed.random_protein(1301, 200)
    parts = [np.zeros(0, np.uint8), q[:1], q[:2], q[:3], q[10:80], q.copy(),
             np.full(500, 1, np.uint8), q[::-1].copy(), hbseed.random_protein(1302, 5000)]
    res = np.concatenate(parts)
    offs = np.zeros(len(parts) + 1, np.uint64)
    offs[1:] = np.cumsum([len(p) for p in parts])
    yield "edges_q200", hbseed.blastp_query(q), res, offs


def main():
    manifest = {}
    for name, spec, res, offs in cases():
        cnt, hsps = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=4)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, seq=spec.seq, ctx_off=spec.ctx_off, ctx_len=spec.ctx_len,
                            xdrop=spec.xdrop, cutoff=spec.cutoff, residues=res, offsets=offs,
                            hsps=hsps.view(np.int32).reshape(-1, 8), counters=cnt)
        manifest[name + ".npz"] = {
            "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest(),
            "hsps": int(len(hsps)), "subjects": int(len(offs) - 1)}
        print(name, len(hsps), "HSPs")
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
