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
                ("window", C.c_int32), ("matrix", C.c_void_p)]


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
                 matrix=None):
        self.arrays = [np.ascontiguousarray(seq, np.uint8),
                       np.ascontiguousarray(ctx_off, np.int32),
                       np.ascontiguousarray(ctx_len, np.int32),
                       np.ascontiguousarray(xdrop, np.int32),
                       np.ascontiguousarray(cutoff, np.int32)]
        self.matrix = None if matrix is None else np.ascontiguousarray(matrix, np.int8)
        s, o, l, x, c = self.arrays
        self.q = _Query(s.ctypes.data, len(s), len(o), o.ctypes.data, l.ctypes.data,
                        x.ctypes.data, c.ctypes.data, 3, threshold, window,
                        None if self.matrix is None else self.matrix.ctypes.data)
        self.lut = _Lut()
        rc = lib().hbo_lut_build(C.byref(self.q), C.byref(self.lut))
        if rc:
            raise RuntimeError(f"hbo_lut_build failed: {rc}")

    @classmethod
    def from_spec(cls, spec, **kw):
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
