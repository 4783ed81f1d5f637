"""FASTA ingestion (hb_fasta_parse; CPU only, no GPU).

Round trips: sequences written as FASTA with wrapped, mixed-case, CRLF and
numbered lines parse back to the same codes and titles; errors name the
offending text offset.
"""
import numpy as np
import pytest

import hbseed


def write_fasta(seqs, titles, rng):
    out = []
    for t, s in zip(titles, seqs):
        out.append(">" + t + ("\r\n" if rng.random() < 0.3 else "\n"))
        txt = hbseed.decode_protein(s)
        w = int(rng.integers(5, 80))
        for i in range(0, len(txt), w):
            line = txt[i:i + w]
            if rng.random() < 0.3:
                line = line.lower()
            if rng.random() < 0.2:
                line = f"{i + 1:>6} " + " ".join(line[j:j + 10] for j in range(0, len(line), 10))
            out.append(line + ("\r\n" if rng.random() < 0.3 else "\n"))
        if rng.random() < 0.2:
            out.append("\n")
    return "".join(out)


@pytest.mark.parametrize("seed", range(5))
def test_round_trip(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 40))
    seqs = [hbseed.random_protein(1000 * seed + i, int(rng.integers(0, 400))) for i in range(n)]
    titles = [f"sp|P{seed}{i:05d}|NAME_{i} protein {i} OS=Synthetic" for i in range(n)]
    text = write_fasta(seqs, titles, rng)
    res, offs, got_titles = hbseed.parse_fasta(text)
    assert len(offs) == n + 1 and got_titles == titles
    for i, s in enumerate(seqs):
        assert np.array_equal(res[offs[i]:offs[i + 1]], s)


def test_stop_gap_and_comment_lines():
    res, offs, titles = hbseed.parse_fasta("; old-style comment\n>a\nAC*\n;x\nD-E\n>b\n")
    assert hbseed.decode_protein(res) == "AC*D-E" and list(offs) == [0, 6, 6]
    assert titles == ["a", "b"]


def test_empty_text_and_empty_records():
    res, offs, titles = hbseed.parse_fasta("")
    assert len(res) == 0 and list(offs) == [0] and titles == []
    res, offs, titles = hbseed.parse_fasta(">x\n>y\n\n>z")
    assert list(offs) == [0, 0, 0, 0] and titles == ["x", "y", "z"]


@pytest.mark.parametrize("text,where", [(">a\nACDX\nAC#D\n", 10), ("ACD\n>a\nAC\n", 0),
                                        (">a\nAC.D\n", 5)])
def test_bad_characters_name_their_offset(text, where):
    with pytest.raises(hbseed.HBError) as e:
        hbseed.parse_fasta(text)
    assert e.value.code == -5 and f"offset {where}" in str(e.value)


def test_capacity_is_reported():
    import ctypes as C
    text = b">a\nACDE\n>b\nFG\n"
    nr, ns = C.c_uint64(), C.c_uint32()
    res = np.zeros(3, np.uint8)
    offs = np.zeros(3, np.uint64)
    rc = hbseed._lib().hb_fasta_parse(text, len(text), res.ctypes.data, len(res),
                                      offs.ctypes.data, None, None, 2, C.byref(nr), C.byref(ns))
    assert rc == -4 and nr.value == 6 and ns.value == 2
