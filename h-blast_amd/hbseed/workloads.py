"""The benchmark workloads as seeded synthetic data (shared by bench.py and
the full-size GPU tests).

* plain BLASTP: a 1,000-residue query (background frequencies, seed 1101)
  against 2,000,000 subjects (log-normal lengths, median 300, sigma 0.6,
  clipped 30-5,000, 5% planted homologs of the query; seed 1100);
* plain BLASTX: a 1,500-nt query (seed 1200) with two coding regions, one
  per strand, that code 70%-identity homologs (one indel) of 100-residue
  segments of two database subjects;
* batched: 2,048 BLASTP queries with log-normal lengths (median 350, sigma
  0.7, clipped 50-3,000; seed 1300) against the plain database;
* served: 20,000,000 subjects (seed 1400, otherwise as the plain database)
  and a stream of 80% BLASTP queries (lengths as the batched ones) and 20%
  BLASTX queries of U[600, 3000] nt, with exponential gaps for Poisson
  arrivals.
"""
import numpy as np

from . import (blastp_query, blastx_nt, blastx_query, homolog, lognormal_lengths, make_db,
               random_protein)


def plain_db(n_subjects=2_000_000):
    """The plain BLASTP query and its database: (query, residues, offsets)."""
    q = random_protein(1101, 1000)
    res, offs = make_db(1100, n_subjects, plant_query=q, plant_frac=0.05)
    return q, res, offs


def blastx_c2(res, offs):
    """The 1,500-nt BLASTX query against the plain database (its coding
    regions are homologs of segments of subjects 123,457 and 1,500,001, or of
    the middle and last subjects of a smaller database)."""
    n = len(offs) - 1
    oids = (123_457, 1_500_001) if n > 1_500_001 else (n // 2, n - 1)
    segs = []
    for k, oid in enumerate(oids):
        a, b = int(offs[oid]), int(offs[oid + 1])
        src = res[a:a + min(100, b - a)]
        segs.append(homolog(1201 + k, src, 0.7, 1))
    return blastx_query(blastx_nt(1200, 1500, segs[0], segs[1]))


def batched_queries(n_queries=2048):
    """The batched workload's queries (residue arrays)."""
    lens = lognormal_lengths(1300, n_queries, 350.0, 0.7, 50, 3000)
    return [random_protein(1300 * 100003 + i, int(L)) for i, L in enumerate(lens)]


def served_db(n_subjects=20_000_000):
    """The served database: (residues, offsets)."""
    q1 = random_protein(1101, 1000)
    return make_db(1400, n_subjects, plant_query=q1, plant_frac=0.05)


def served_stream(n, seed=1400):
    """Arrival gaps (exponential, mean 1: divide by the rate in queries/s)
    and the query mix [("blastp" | "blastx", QuerySpec)]."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0, n)
    kinds = rng.random(n) < 0.8
    lens = lognormal_lengths(seed, n, 350.0, 0.7, 50, 3000)
    nts = rng.integers(600, 3001, n)
    src = random_protein(1101, 1000)
    specs = []
    for i in range(n):
        if kinds[i]:
            specs.append(("blastp", blastp_query(random_protein(seed * 100003 + i, int(lens[i])))))
        else:
            seg_len = int(min(100, nts[i] // 7))
            a = int(rng.integers(0, 1000 - seg_len))
            seg = homolog(seed * 7 + i, src[a:a + seg_len], 0.7, 1)
            nt = blastx_nt(seed * 13 + i, int(nts[i]), seg, seg[::-1].copy())
            specs.append(("blastx", blastx_query(nt)))
    return gaps, specs
