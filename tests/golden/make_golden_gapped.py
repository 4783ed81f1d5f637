#!/usr/bin/env python3
"""Writes the golden fixtures of the preliminary gapped extension.

For each seed-stage fixture (tests/golden/<name>.npz, written by
make_golden.py) this holds the oracle's gapped output for that fixture's
HSPs: gapped records in (oid, save order) as int32[n, 10], per-subject
counts (saved, extensions), and the parameters (gap open 11, extend 1,
X-drop 38, cutoff per context), as .npz without pickles, with a SHA-256
manifest.  The oracle's gapped stage is pinned by
tests/test_oracle_gapped.py.

Run from the repository root after building:
python3 tests/golden/make_golden_gapped.py
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

SOURCES = {"blastp_q300_s2000.npz": 45, "blastx_nt600_s1000.npz": 40, "edges_q200.npz": 40}
GAP = (11, 1, 38)


def main():
    manifest = {}
    for src, cut in sorted(SOURCES.items()):
        z = np.load(os.path.join(HERE, src))
        spec = hbseed.QuerySpec(z["seq"], z["ctx_off"], z["ctx_len"], z["xdrop"], z["cutoff"])
        hsps = z["hsps"].copy().view(oracle_lib.HSP_DTYPE).reshape(-1)
        cutoffs = np.full(len(spec.ctx_off), cut, np.int32)
        oq = oracle_lib.OracleQuery.from_spec(spec)
        cnt, g = oracle_lib.gapped_db(oq, z["residues"], z["offsets"], hsps, cutoffs, *GAP,
                                      threads=4)
        name = "gapped_" + src
        path = os.path.join(HERE, name)
        np.savez_compressed(path, cutoffs=cutoffs, gap=np.array(GAP, np.int32),
                            gapped=g.view(np.int32).reshape(-1, 10), counts=cnt)
        manifest[name] = {"source": src,
                          "sha256": hashlib.sha256(open(path, "rb").read()).hexdigest(),
                          "gapped": int(len(g)), "extensions": int(cnt[:, 1].sum())}
        print(name, len(g), "gapped HSPs")
    with open(os.path.join(HERE, "MANIFEST_GAPPED.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
