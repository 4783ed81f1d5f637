"""ctypes view of the CPU oracle (oracle/libhbo.so) -- test infrastructure.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
this module, and only as the checker or the timed CPU baseline.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "libhbo.so")

HSP_DTYPE = np.dtype([("oid", "<u4"), ("context", "<i4"), ("q_start", "<i4"),
                      ("s_start", "<i4"), ("q_off", "<i4"), ("s_off", "<i4"),
                      ("len", "<i4"), ("score", "<i4")])


class _Query(C.Structure):
    _fields_ = [("seq", C.c_void_p), ("len", C.c_int32), ("n_ctx", C.c_int32),
                ("ctx_off", C.c_void_p), ("ctx_len", C.c_void_p), ("xdrop", C.c_void_p),
                ("cutoff", C.c_void_p), ("word_size", C.c_int32), ("threshold", C.c_int32),
                ("window", C.c_int32), ("matrix", C.c_void_p), ("lut_mask", C.c_void_p)]


class _Lut(C.Structure):
    _fields_ = [("start", C.POINTER(C.c_uint32)), ("entries", C.POINTER(C.c_uint32)),
                ("n_entries", C.c_uint32)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        L = C.CDLL(LIB_PATH)
        L.hbo_blosum62.restype = C.c_void_p
        L.hbo_lut_build.argtypes = [C.POINTER(_Query), C.POINTER(_Lut)]
        L.hbo_lut_free.argtypes = [C.POINTER(_Lut)]
        L.hbo_scan_db.restype = C.c_int64
        L.hbo_scan_db.argtypes = [C.POINTER(_Query), C.POINTER(_Lut), C.c_void_p, C.c_void_p,
                                  C.c_uint32, C.c_int, C.POINTER(C.c_void_p), C.c_void_p]
        L.hbo_free.argtypes = [C.c_void_p]
        i32 = C.c_int32
        pi = C.POINTER(C.c_int32)
        L.hbo_extend_left.restype = i32
        L.hbo_extend_left.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, i32, i32, i32, i32, pi]
        L.hbo_extend_right.restype = i32
        L.hbo_extend_right.argtypes = [C.c_void_p, C.c_void_p, i32, C.c_void_p, i32, i32, i32,
                                       i32, i32, pi, pi]
        L.hbo_extend_two_hit.restype = i32
        L.hbo_extend_two_hit.argtypes = [C.c_void_p, C.c_void_p, i32, C.c_void_p, i32, i32, i32,
                                         i32, i32, pi, pi, pi, pi, pi]
        _L = L
    return _L


def blosum62():
    p = lib().hbo_blosum62()
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int8)), shape=(28, 28)).copy()


class OracleQuery:
    def __init__(self, seq, ctx_off, ctx_len, xdrop, cutoff, threshold=11, window=40,
                 matrix=None, lut_mask=None):
        self.arrays = [np.ascontiguousarray(seq, np.uint8),
                       np.ascontiguousarray(ctx_off, np.int32),
                       np.ascontiguousarray(ctx_len, np.int32),
                       np.ascontiguousarray(xdrop, np.int32),
                       np.ascontiguousarray(cutoff, np.int32)]
        self.matrix = None if matrix is None else np.ascontiguousarray(matrix, np.int8)
        self.lut_mask = None if lut_mask is None else np.ascontiguousarray(lut_mask, np.uint8)
        s, o, l, x, c = self.arrays
        self.q = _Query(s.ctypes.data, len(s), len(o), o.ctypes.data, l.ctypes.data,
                        x.ctypes.data, c.ctypes.data, 3, threshold, window,
                        None if self.matrix is None else self.matrix.ctypes.data,
                        None if self.lut_mask is None else self.lut_mask.ctypes.data)
        self.lut = _Lut()
        rc = lib().hbo_lut_build(C.byref(self.q), C.byref(self.lut))
        if rc:
            raise RuntimeError(f"hbo_lut_build failed: {rc}")

    @classmethod
    def from_spec(cls, spec, **kw):
        if getattr(spec, "lut_mask", None) is not None:
            kw.setdefault("lut_mask", spec.lut_mask)
        return cls(spec.seq, spec.ctx_off, spec.ctx_len, spec.xdrop, spec.cutoff, **kw)

    def lut_arrays(self):
        start = np.ctypeslib.as_array(self.lut.start, shape=(32769,)).copy()
        n = self.lut.n_entries
        ent = np.ctypeslib.as_array(self.lut.entries, shape=(max(n, 1),))[:n].copy()
        return start, ent

    def scan(self, residues, offsets, threads=1):
        residues = np.ascontiguousarray(residues, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        n = len(offsets) - 1
        out = C.c_void_p()
        counters = np.zeros(3 * max(n, 1), np.uint32)
        k = lib().hbo_scan_db(C.byref(self.q), C.byref(self.lut), residues.ctypes.data,
                              offsets.ctypes.data, n, threads, C.byref(out),
                              counters.ctypes.data)
        if k < 0:
            raise RuntimeError(f"hbo_scan_db failed: {k}")
        hsps = np.zeros(k, HSP_DTYPE)
        if k:
            C.memmove(hsps.ctypes.data, out.value, k * HSP_DTYPE.itemsize)
        lib().hbo_free(out)
        return counters[:3 * n].reshape(n, 3), hsps

    def check_hsps(self, residues, offsets, hsps):
        """Property check of device HSP records (hbo_hsp_check): returns
        (number of bad records, index of the first)."""
        residues = np.ascontiguousarray(residues, np.uint8)
        offsets = np.ascontiguousarray(offsets, np.uint64)
        h = np.ascontiguousarray(hsps).view(HSP_DTYPE)
        first = C.c_uint64()
        L = lib()
        L.hbo_hsp_check.restype = C.c_uint64
        L.hbo_hsp_check.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                    C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)]
        bad = L.hbo_hsp_check(C.addressof(self.q), residues.ctypes.data, offsets.ctypes.data,
                              len(offsets) - 1, h.ctypes.data, len(h), C.byref(first))
        return int(bad), int(first.value)

    def __del__(self):
        try:
            lib().hbo_lut_free(C.byref(self.lut))
        except Exception:
            pass


def extend_left(m, s, q, s_off, q_off, x, maxscore=0):
    s = np.ascontiguousarray(s, np.uint8)
    q = np.ascontiguousarray(q, np.uint8)
    m = np.ascontiguousarray(m, np.int8)
    ln = C.c_int32()
    sc = lib().hbo_extend_left(m.ctypes.data, s.ctypes.data, q.ctypes.data, s_off, q_off, x,
                               maxscore, C.byref(ln))
    return sc, ln.value


def extend_right(m, s, q, s_off, q_off, x, maxscore):
    s = np.ascontiguousarray(s, np.uint8)
    q = np.ascontiguousarray(q, np.uint8)
    m = np.ascontiguousarray(m, np.int8)
    ln, last = C.c_int32(), C.c_int32()
    sc = lib().hbo_extend_right(m.ctypes.data, s.ctypes.data, len(s), q.ctypes.data, len(q),
                                s_off, q_off, x, maxscore, C.byref(ln), C.byref(last))
    return sc, ln.value, last.value


def extend_two_hit(m, s, q, s_left_off, s_right_off, q_right_off, x):
    s = np.ascontiguousarray(s, np.uint8)
    q = np.ascontiguousarray(q, np.uint8)
    m = np.ascontiguousarray(m, np.int8)
    hq, hs, hl, re, sl = (C.c_int32() for _ in range(5))
    sc = lib().hbo_extend_two_hit(m.ctypes.data, s.ctypes.data, len(s), q.ctypes.data, len(q),
                                  s_left_off, s_right_off, q_right_off, x, C.byref(hq),
                                  C.byref(hs), C.byref(hl), C.byref(re), C.byref(sl))
    return dict(score=sc, q_start=hq.value, s_start=hs.value, len=hl.value,
                right=re.value, s_last=sl.value)


# ------------------------------------------------------------ gapped extension
GHSP_DTYPE = np.dtype([("oid", "<u4"), ("context", "<i4"), ("q_start", "<i4"),
                       ("q_end", "<i4"), ("s_start", "<i4"), ("s_end", "<i4"),
                       ("q_gstart", "<i4"), ("s_gstart", "<i4"), ("score", "<i4"),
                       ("pad", "<i4")])


class _GapParams(C.Structure):
    _fields_ = [("gap_open", C.c_int32), ("gap_extend", C.c_int32), ("x_dropoff", C.c_int32)]


def _gap_lib():
    L = lib()
    if not getattr(L, "_gap_ready", False):
        i32, pi, vp = C.c_int32, C.POINTER(C.c_int32), C.c_void_p
        L.hbo_semigapped.restype = i32
        L.hbo_semigapped.argtypes = [vp, vp, vp, i32, i32, C.c_int, i32, i32, i32, pi, pi]
        L.hbo_gapped_start.restype = i32
        L.hbo_gapped_start.argtypes = [vp, vp, vp, i32, i32, i32, i32]
        L.hbo_prot_gapped.restype = i32
        L.hbo_prot_gapped.argtypes = [vp, vp, i32, vp, i32, i32, i32, C.POINTER(_GapParams), vp]
        L.hbo_gapped_db.restype = C.c_int64
        L.hbo_gapped_db.argtypes = [C.POINTER(_Query), vp, vp, C.c_uint32, vp, C.c_uint64,
                                    C.POINTER(_GapParams), vp, C.c_int, C.POINTER(vp), vp]
        L._gap_ready = True
    return L


def semigapped(m, A, B, M, N, reverse, gap_open=11, gap_extend=1, x_dropoff=38):
    """Score-only X-drop DP: returns (score, a_off, b_off)."""
    A = np.ascontiguousarray(A, np.uint8)
    B = np.ascontiguousarray(B, np.uint8)
    m = np.ascontiguousarray(m, np.int8)
    a, b = C.c_int32(), C.c_int32()
    sc = _gap_lib().hbo_semigapped(m.ctypes.data, A.ctypes.data, B.ctypes.data, M, N,
                                   int(reverse), gap_open, gap_extend, x_dropoff,
                                   C.byref(a), C.byref(b))
    return sc, a.value, b.value


def gapped_start(m, q, s, q_start, q_len, s_start, s_len):
    q = np.ascontiguousarray(q, np.uint8)
    s = np.ascontiguousarray(s, np.uint8)
    m = np.ascontiguousarray(m, np.int8)
    return _gap_lib().hbo_gapped_start(m.ctypes.data, q.ctypes.data, s.ctypes.data, q_start,
                                       q_len, s_start, s_len)


def prot_gapped(m, q, s, q_off, s_off, gap_open=11, gap_extend=1, x_dropoff=38):
    """Two-sided extension: returns (score, q_start, q_end, s_start, s_end)."""
    q = np.ascontiguousarray(q, np.uint8)
    s = np.ascontiguousarray(s, np.uint8)
    m = np.ascontiguousarray(m, np.int8)
    co = np.zeros(4, np.int32)
    gp = _GapParams(gap_open, gap_extend, x_dropoff)
    sc = _gap_lib().hbo_prot_gapped(m.ctypes.data, q.ctypes.data, len(q), s.ctypes.data,
                                    len(s), q_off, s_off, C.byref(gp), co.ctypes.data)
    return (sc, *map(int, co))


def gapped_db(oq, residues, offsets, hsps, cutoffs, gap_open=11, gap_extend=1, x_dropoff=38,
              threads=1):
    """Preliminary gapped extension of every subject's ungapped HSPs (grouped by
    ascending oid).  Returns (counts[n, 2] = saved, extensions; gapped records)."""
    residues = np.ascontiguousarray(residues, np.uint8)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    hsps = np.ascontiguousarray(hsps, HSP_DTYPE)
    cut = np.ascontiguousarray(cutoffs, np.int32)
    n = len(offsets) - 1
    counts = np.zeros(2 * max(n, 1), np.int32)
    out = C.c_void_p()
    gp = _GapParams(gap_open, gap_extend, x_dropoff)
    k = _gap_lib().hbo_gapped_db(C.byref(oq.q), residues.ctypes.data, offsets.ctypes.data, n,
                                 hsps.ctypes.data, len(hsps), C.byref(gp), cut.ctypes.data,
                                 threads, C.byref(out), counts.ctypes.data)
    if k < 0:
        raise RuntimeError(f"hbo_gapped_db failed: {k}")
    recs = np.zeros(k, GHSP_DTYPE)
    if k:
        C.memmove(recs.ctypes.data, out.value, k * GHSP_DTYPE.itemsize)
    lib().hbo_free(out)
    return counts[:2 * n].reshape(n, 2), recs


class _HdbVolume(C.Structure):
    _fields_ = [("name", C.c_char_p), ("bytes", C.c_uint64), ("volume_seqs", C.c_int32),
                ("gpu_seqs", C.c_int32), ("store_limit", C.c_int32),
                ("break_limit", C.c_int32), ("n_groups", C.c_int32),
                ("group_len", C.c_void_p), ("cpu_rest_bytes", C.c_uint64)]


def _hdb_lib():
    L = lib()
    if not getattr(L, "_hdb_ready", False):
        i32, vp, u64 = C.c_int32, C.c_void_p, C.c_uint64
        L.hbo_hdb_gpu_seqs.restype = i32
        L.hbo_hdb_gpu_seqs.argtypes = [i32]
        L.hbo_hdb_layout.restype = C.c_int
        L.hbo_hdb_layout.argtypes = [vp, i32, i32, i32, vp, C.POINTER(u64)]
        L.hbo_hdb_pack.restype = C.c_int
        L.hbo_hdb_pack.argtypes = [vp, vp, i32, i32, i32, vp, vp, u64]
        L.hbo_hdb_unpack.restype = C.c_int
        L.hbo_hdb_unpack.argtypes = [vp, u64, i32, i32, i32, vp, vp, vp]
        L.hbo_hdbi_format.restype = C.c_int64
        L.hbo_hdbi_format.argtypes = [i32, i32, i32, i32, vp, C.c_char_p, u64]
        L._hdb_ready = True
    return L


def hdb_gpu_seqs(n):
    return _hdb_lib().hbo_hdb_gpu_seqs(int(n))


def hdb_layout(offsets, n_seqs, blocksx, threadsx):
    offsets = np.ascontiguousarray(offsets, np.uint64)
    G = blocksx * threadsx
    gl = np.zeros(max((n_seqs + G - 1) // G, 1), np.int32)
    nb = C.c_uint64()
    rc = _hdb_lib().hbo_hdb_layout(offsets.ctypes.data, n_seqs, blocksx, threadsx,
                                   gl.ctypes.data, C.byref(nb))
    if rc:
        raise RuntimeError(f"hbo_hdb_layout failed: {rc}")
    return gl[:(n_seqs + G - 1) // G].copy(), nb.value


def hdb_pack(residues, offsets, n_seqs, blocksx, threadsx):
    residues = np.ascontiguousarray(residues, np.uint8)
    offsets = np.ascontiguousarray(offsets, np.uint64)
    gl, nb = hdb_layout(offsets, n_seqs, blocksx, threadsx)
    out = np.zeros(max(nb, 1), np.uint8)
    rc = _hdb_lib().hbo_hdb_pack(residues.ctypes.data, offsets.ctypes.data, n_seqs, blocksx,
                                 threadsx, gl.ctypes.data, out.ctypes.data, nb)
    if rc:
        raise RuntimeError(f"hbo_hdb_pack failed: {rc}")
    return out[:nb], gl


def hdb_unpack(data, n_seqs, blocksx, threadsx, group_len):
    """The reference kernel's walk over a .hDb -> (residues, offsets)."""
    data = np.ascontiguousarray(data, np.uint8)
    if data.size % 4:
        data = np.concatenate([data, np.zeros(4 - data.size % 4, np.uint8)])
    gl = np.ascontiguousarray(group_len, np.int32)
    offs = np.zeros(n_seqs + 1, np.uint64)
    L = _hdb_lib()
    rc = L.hbo_hdb_unpack(data.ctypes.data, data.size, n_seqs, blocksx, threadsx,
                          gl.ctypes.data, None, offs.ctypes.data)
    if rc:
        raise RuntimeError(f"hbo_hdb_unpack failed: {rc}")
    res = np.zeros(max(int(offs[-1]), 1), np.uint8)
    rc = L.hbo_hdb_unpack(data.ctypes.data, data.size, n_seqs, blocksx, threadsx,
                          gl.ctypes.data, res.ctypes.data, offs.ctypes.data)
    if rc:
        raise RuntimeError(f"hbo_hdb_unpack failed: {rc}")
    return res[:int(offs[-1])].copy(), offs


def hdbi_format(blocksx, threadsx, volumes, percentage=99):
    arr = (_HdbVolume * max(len(volumes), 1))()
    keep = []
    for i, v in enumerate(volumes):
        gl = np.ascontiguousarray(v["group_len"], np.int32)
        keep.append(gl)
        arr[i] = _HdbVolume(v["name"].encode(), v["bytes"], v["volume_seqs"], v["gpu_seqs"],
                            v["store_limit"], v["break_limit"], gl.size, gl.ctypes.data,
                            v["cpu_rest_bytes"])
    L = _hdb_lib()
    n = L.hbo_hdbi_format(blocksx, threadsx, percentage, len(volumes), C.addressof(arr), None, 0)
    buf = C.create_string_buffer(n + 1)
    L.hbo_hdbi_format(blocksx, threadsx, percentage, len(volumes), C.addressof(arr), buf, n)
    return buf.raw[:n].decode()


# --- CPU/GPU split of a volume (hbo_balance.c) ---------------------------

class _Options(C.Structure):
    _fields_ = [("gpu_ratio", C.c_float), ("fixed_ratio", C.c_int32),
                ("capability_rate", C.c_double), ("max_hits", C.c_int32),
                ("mem_ecc", C.c_int32), ("load_balancer", C.c_int32),
                ("load_factor", C.c_int32), ("subject_chunk", C.c_int32),
                ("base_value", C.c_int32)]


_OPTION_KEYS = ["gpu_ratio", "fixed_ratio", "capability_rate", "max_hits_per_subject", "mem_ecc",
                "load_balancer", "load_factor", "subject_chunk", "base_value"]


def _balance_lib():
    L = lib()
    if not getattr(L, "_balance_ready", False):
        i32, u32, vp, u64 = C.c_int32, C.c_uint32, C.c_void_p, C.c_uint64
        L.hbo_options_default.argtypes = [vp]
        L.hbo_options_default.restype = None
        L.hbo_options_parse.argtypes = [C.c_char_p, u64, vp]
        L.hbo_options_parse.restype = C.c_int
        L.hbo_split.argtypes = [vp, i32, i32, u64, i32, i32, i32, C.POINTER(C.c_float),
                                C.POINTER(i32), C.POINTER(i32), C.POINTER(i32), C.POINTER(u64)]
        L.hbo_split.restype = C.c_int
        L.hbo_next_ratio.argtypes = [vp, u32, C.c_float, u32, u32, vp, vp, vp, C.c_int, C.c_int,
                                     u32, C.c_double, C.c_double, C.POINTER(C.c_float)]
        L.hbo_next_ratio.restype = C.c_int
        L._balance_ready = True
    return L


def runtime_options(text=None):
    """Options dict (keys as the product's), defaults or parsed; None when
    the oracle rejects the text."""
    o = _Options()
    L = _balance_lib()
    if text is None:
        L.hbo_options_default(C.byref(o))
    else:
        if isinstance(text, str):
            text = text.encode()
        if L.hbo_options_parse(text, len(text), C.byref(o)) != 0:
            return None
    return dict(zip(_OPTION_KEYS, (getattr(o, f[0]) for f in _Options._fields_)))


def split(group_len, gpu_seqs, nbytes, blocksx, threadsx, cpu_threads, ratio=-1.0, fixed=1):
    gl = np.ascontiguousarray(group_len, np.int32)
    r, f = C.c_float(ratio), C.c_int32(fixed)
    s, g, b = C.c_int32(), C.c_int32(), C.c_uint64()
    rc = _balance_lib().hbo_split(gl.ctypes.data if gl.size else None, gl.size, int(gpu_seqs),
                                  int(nbytes), threadsx, blocksx * threadsx, cpu_threads,
                                  C.byref(r), C.byref(f), C.byref(s), C.byref(g), C.byref(b))
    if rc != 0:
        raise RuntimeError(f"hbo_split failed: {rc}")
    return dict(gpu_seqs=s.value, n_groups=g.value, bytes=b.value), r.value, f.value


def next_ratio(options, volume, ratio, cpu_threads, volume_bytes, ratio_ring, hsp_ring,
               gpu_seconds, cpu_seconds, gapped=False, query_length=0, avg_query_length=0,
               gpu_cards=1):
    """-> new ratio, or None where the oracle reports an error."""
    o = _Options(*(options[k] for k in _OPTION_KEYS))
    vb = np.ascontiguousarray(volume_bytes, np.uint64)
    rr = np.ascontiguousarray(ratio_ring, np.float32)
    hr = np.ascontiguousarray(hsp_ring, np.int64)
    out = C.c_float()
    rc = _balance_lib().hbo_next_ratio(C.byref(o), volume, ratio, cpu_threads, gpu_cards,
                                       vb.ctypes.data, rr.ctypes.data, hr.ctypes.data,
                                       int(bool(gapped)), query_length, avg_query_length,
                                       gpu_seconds, cpu_seconds, C.byref(out))
    return None if rc != 0 else out.value


# --- SEG low-complexity masking (hbo_seg.c) --------------------------------

def _seg_lib():
    L = lib()
    if not getattr(L, "_seg_ready", False):
        i32, vp = C.c_int32, C.c_void_p
        L.hbo_seg_entropy.argtypes = [vp]
        L.hbo_seg_entropy.restype = C.c_double
        L.hbo_seg_lnprob.argtypes = [vp, i32]
        L.hbo_seg_lnprob.restype = C.c_double
        L.hbo_seg_mask.argtypes = [vp, i32, i32, vp, vp, i32, C.c_double, C.c_double, i32, i32,
                                   vp]
        L.hbo_seg_mask.restype = C.c_int
        L._seg_ready = True
    return L


def seg_entropy(counts28):
    c = np.ascontiguousarray(counts28, np.int32)
    return _seg_lib().hbo_seg_entropy(c.ctypes.data)


def seg_lnprob(counts28, n):
    c = np.ascontiguousarray(counts28, np.int32)
    return _seg_lib().hbo_seg_lnprob(c.ctypes.data, int(n))


def seg_mask(seq, ctx_off, ctx_len, window=12, locut=2.2, hicut=2.5, maxbogus=2, maxtrim=100):
    seq = np.ascontiguousarray(seq, np.uint8)
    off = np.ascontiguousarray(ctx_off, np.int32)
    ln = np.ascontiguousarray(ctx_len, np.int32)
    out = np.zeros(seq.size, np.uint8)
    rc = _seg_lib().hbo_seg_mask(seq.ctypes.data, seq.size, off.size, off.ctypes.data,
                                 ln.ctypes.data, window, locut, hicut, maxbogus, maxtrim,
                                 out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"hbo_seg_mask failed: {rc}")
    return out
