"""The C ABI library: loads, exports every function include/hb_seed.h
declares, and rejects bad arguments before touching a device (CPU only)."""
import ctypes as C
import os
import re

import numpy as np

import hbseed

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hb_seed.h")


def declared():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    return sorted(set(re.findall(r"\b(hb_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(hbseed.lib_path())
    names = declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hbseed.EXPORTS) == names


def test_error_strings():
    L = hbseed._lib()
    for code in range(0, -9, -1):
        assert L.hb_strerror(code)
    assert L.hb_strerror(-4) == b"output array too small"


def test_argument_checks_need_no_device():
    L = hbseed._lib()
    h = C.c_void_p()
    assert L.hb_create(0, None) == -1
    assert L.hb_db_upload(None, None, None, 0, C.byref(h)) == -1
    assert L.hb_query_prepare(None, None, 0, 0, None, None, None, None, None, C.byref(h)) == -1
    assert L.hb_scan_start(None, None, None, C.byref(h)) == -1
    assert L.hb_scan_results(None, None, None, 0, None) == -1
    assert L.hb_query_prepare_masked(None, None, 0, 0, None, None, None, None, None, None,
                                     C.byref(h)) == -1
    gp = (C.c_int32 * 4)(11, 1, 38, 0)
    cut = (C.c_int32 * 1)(50)
    assert L.hb_scan_gapped(None, C.byref(gp), cut, None, None, 0, None) == -1
    assert L.hb_gapped_search(None, None, None, 0, 0, None, None, None, None, 0, C.byref(gp),
                              cut, None, None, 0, None) == -1
    bad = (C.c_int32 * 4)(-1, 1, 38, 0)
    assert L.hb_gapped_search(None, None, None, 0, 0, None, None, None, None, 0, C.byref(bad),
                              cut, None, None, 0, None) == -1
    assert L.hb_batch_add_masked(None, None, 0, 0, None, None, None, None, None, None) == -1
    rc = L.hb_create(0, C.byref(h))
    if rc == 0:      # a GPU is present: release it
        L.hb_destroy(h)
    else:
        assert rc in (-1, -3)


def test_encode_and_translate_on_host():
    assert list(hbseed.encode_protein("ACDW*")) == [1, 3, 4, 20, 25]
    try:
        hbseed.encode_protein("AB1")
        raise AssertionError("expected HB_ERR_RESIDUE")
    except hbseed.HBError as e:
        assert e.code == -5
    spec = hbseed.blastx_query("ATGGCC" * 5 + "TAA")
    assert hbseed.decode_protein(spec.seq[:11]) == "MAMAMAMAMA*"
    assert list(spec.ctx_len) == [11, 10, 10, 11, 10, 10]
    assert len(spec.seq) == hbseed._lib().hb_blastx_concat_len(33)


def test_synthetic_generator_properties():
    lens = hbseed.lognormal_lengths(5, 20000, 300.0, 0.6, 30, 5000)
    assert lens.min() >= 30 and lens.max() <= 5000
    assert abs(np.median(lens) - 300) < 10
    res, offs = hbseed.make_db(5, 2000)
    assert np.array_equal(np.diff(offs), hbseed.lognormal_lengths(5, 2000, 300.0, 0.6, 30, 5000))
    freq = np.bincount(res, minlength=28) / len(res)
    assert abs(freq[11] - 0.0902) < 0.003     # L, Robinson & Robinson 90.19 per 1000
    assert abs(freq[20] - 0.0133) < 0.002     # W
    assert freq[[0, 2, 21, 23, 24, 25, 26, 27]].sum() == 0
    r2, o2 = hbseed.make_db(5, 2000)
    assert np.array_equal(res, r2)


def test_result_buffers_grow_and_hand_back_views():
    """ResultBuffers: an array is allocated on the first request, handed back
    as a view of its first n elements while it is large enough, and replaced
    by a larger one (1.25x the request) when it is not."""
    b = hbseed.ResultBuffers()
    a = b.get("hsps", hbseed.HSP_DTYPE, 100)
    assert len(a) == 100 and a.base is not None and len(a.base) == 125
    a[:] = 7
    a2 = b.get("hsps", hbseed.HSP_DTYPE, 40)
    assert len(a2) == 40 and a2.base is a.base and (a2.view(np.int32) == 7).all()
    a3 = b.get("hsps", hbseed.HSP_DTYPE, 200)
    assert len(a3) == 200 and len(a3.base) == 250 and a3.base is not a.base
    assert len(b.get("hsps", hbseed.HSP_DTYPE, 0)) == 0
    s = b.get("stats", hbseed.STATS_DTYPE, 10)
    assert s.dtype == hbseed.STATS_DTYPE and len(s) == 10
