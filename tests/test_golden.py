"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the files match their manifest, the synthetic generator reproduces the
fixture inputs from their seeds, and the oracle reproduces the fixture
outputs.  GPU: the HIP stage reproduces the fixture outputs bit-exactly.
"""
import hashlib
import json
import os

import numpy as np
import pytest

import hbseed
import oracle_lib

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "MANIFEST.json")))
NAMES = sorted(MANIFEST)


def load(name):
    z = np.load(os.path.join(HERE, name))   # allow_pickle defaults to False
    spec = hbseed.QuerySpec(z["seq"], z["ctx_off"], z["ctx_len"], z["xdrop"], z["cutoff"])
    return spec, z["residues"], z["offsets"], z["hsps"], z["counters"]


@pytest.mark.parametrize("name", NAMES)
def test_manifest_hash(name):
    data = open(os.path.join(HERE, name), "rb").read()
    assert hashlib.sha256(data).hexdigest() == MANIFEST[name]["sha256"]


def test_generator_reproduces_inputs():
    spec, res, offs, _, _ = load("blastp_q300_s2000.npz")
    q = hbseed.random_protein(1001, 300)
    r2, o2 = hbseed.make_db(1000, 2000, plant_query=q, plant_frac=0.1)
    assert np.array_equal(spec.seq, q)
    assert np.array_equal(res, r2) and np.array_equal(offs, o2)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_reproduces_fixture(name):
    spec, res, offs, hsps, cnt = load(name)
    c2, h2 = oracle_lib.OracleQuery.from_spec(spec).scan(res, offs, threads=2)
    assert np.array_equal(h2.view(np.int32).reshape(-1, 8), hsps)
    assert np.array_equal(c2, cnt)


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_reproduces_fixture(name):
    spec, res, offs, hsps, cnt = load(name)
    with hbseed.Context(0) as ctx:
        db = hbseed.Database(ctx, res, offs)
        stats, got = hbseed.search(ctx, db, spec)
        db.close()
    assert np.array_equal(got.view(np.int32).reshape(-1, 8), hsps)
    assert np.array_equal(stats["lookup_hits"], cnt[:, 0])
    assert np.array_equal(stats["extensions"], cnt[:, 1])
    assert np.array_equal(stats["n_hsp"], cnt[:, 2])
