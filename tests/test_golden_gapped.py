"""Golden fixtures of the gapped stage (tests/golden/gapped_*.npz, written by
tests/golden/make_golden_gapped.py from the seed-stage fixtures).

CPU: the files match their manifest and the oracle reproduces them from the
seed-stage fixture's HSPs.  GPU: hb_scan_gapped on the fixture's inputs
reproduces them bit-exactly.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import hbseed
import oracle_lib

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "MANIFEST_GAPPED.json")))
NAMES = sorted(MANIFEST)


def load(name):
    src = np.load(os.path.join(HERE, MANIFEST[name]["source"]))   # no pickles
    g = np.load(os.path.join(HERE, name))
    spec = hbseed.QuerySpec(src["seq"], src["ctx_off"], src["ctx_len"], src["xdrop"],
                            src["cutoff"])
    hsps = src["hsps"].copy().view(oracle_lib.HSP_DTYPE).reshape(-1)
    return spec, src["residues"], src["offsets"], hsps, g


@pytest.mark.parametrize("name", NAMES)
def test_manifest_hash(name):
    data = open(os.path.join(HERE, name), "rb").read()
    assert hashlib.sha256(data).hexdigest() == MANIFEST[name]["sha256"]


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_gapped_fixture(name):
    spec, res, offs, hsps, g = load(name)
    oq = oracle_lib.OracleQuery.from_spec(spec)
    cnt, got = oracle_lib.gapped_db(oq, res, offs, hsps, g["cutoffs"], *g["gap"].tolist(),
                                    threads=2)
    assert np.array_equal(got.view(np.int32).reshape(-1, 10), g["gapped"])
    assert np.array_equal(cnt, g["counts"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_gapped_fixture(name):
    spec, res, offs, _, g = load(name)
    gap_open, gap_extend, x = g["gap"].tolist()
    with hbseed.Context(0) as ctx:
        db = hbseed.Database(ctx, res, offs)
        q = hbseed.Query(ctx, spec)
        s = hbseed.Scan(ctx, db, q)
        gst, got = s.gapped(g["cutoffs"], gap_open, gap_extend, x)
        s.close()
        q.close()
        db.close()
    assert np.array_equal(got.view(np.int32).reshape(-1, 10), g["gapped"])
    assert np.array_equal(gst["n_hsp"], g["counts"][:, 0])
    assert np.array_equal(gst["extensions"], g["counts"][:, 1])
