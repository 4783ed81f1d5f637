"""hbseed -- Python view of the MI355X seed-extension stage (libhbseed.so).

The stage itself is C++/HIP behind the C ABI in include/hb_seed.h; this
module only loads that library with ctypes and moves numpy arrays across
the boundary (tests and the benchmark use it).  Nothing here computes
results: if the native library is missing, loading raises.

The calls mirror the reference's host interface
(h_blast_v1.1/H-BLAST/ncbi_blast_files/gpu_cpu_common.h:43-126):
``Database`` replaces Read_H_BLAST_GPU_DB, ``Query`` the lookup-table
staging in H_BLAST, ``Scan.start`` H_BLAST_pre and ``Scan.results``
H_BLAST_get_data.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBDIR = os.path.join(os.path.dirname(_HERE), "lib")

HB_OK = 0
HB_ERR_NOMEM = -2
ERRORS = {
    -1: "HB_ERR_ARG", -2: "HB_ERR_NOMEM", -3: "HB_ERR_DEVICE", -4: "HB_ERR_CAPACITY",
    -5: "HB_ERR_RESIDUE", -6: "HB_ERR_STATE", -7: "HB_ERR_UNSUPPORTED", -8: "HB_ERR_BUSY",
    -9: "HB_ERR_FORMAT",
}
DEFAULT_XDROP = 16
DEFAULT_CUTOFF = 41
STATE_AUTO, STATE_LDS, STATE_GLOBAL, STATE_LDS16 = 0, 1, 2, 3
STDAA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"

HSP_DTYPE = np.dtype([("oid", "<u4"), ("context", "<i4"), ("q_start", "<i4"),
                      ("s_start", "<i4"), ("q_off", "<i4"), ("s_off", "<i4"),
                      ("len", "<i4"), ("score", "<i4")])
STATS_DTYPE = np.dtype([("lookup_hits", "<u4"), ("extensions", "<u4"), ("n_hsp", "<u4"),
                        ("first", "<i8")], align=True)
GHSP_DTYPE = np.dtype([("oid", "<u4"), ("context", "<i4"), ("q_start", "<i4"),
                       ("q_end", "<i4"), ("s_start", "<i4"), ("s_end", "<i4"),
                       ("q_gstart", "<i4"), ("s_gstart", "<i4"), ("score", "<i4"),
                       ("pad", "<i4")])
GSTATS_DTYPE = np.dtype([("extensions", "<u4"), ("n_hsp", "<u4"), ("first", "<i8")])
# hb_subject_hits: one entry per subject with HSPs (hb_scan_results_compact)
SUBJ_HITS_DTYPE = np.dtype([("oid", "<u4"), ("n_hsp", "<u4"), ("lookup_hits", "<u4"),
                            ("extensions", "<u4"), ("first", "<u8")])
DEFAULT_GAP_OPEN = 11
DEFAULT_GAP_EXTEND = 1
DEFAULT_GAP_XDROP = 38

# Every symbol include/hb_seed.h declares (checked by the CPU tests).
EXPORTS = [
    "hb_params_default", "hb_strerror", "hb_blosum62", "hb_create", "hb_destroy",
    "hb_mem_info", "hb_synchronize", "hb_host_alloc", "hb_host_free", "hb_db_upload", "hb_db_free", "hb_db_subjects",
    "hb_db_residues", "hb_query_prepare", "hb_query_free", "hb_query_lut_info",
    "hb_query_lut_fetch", "hb_blastx_concat_len", "hb_blastx_translate",
    "hb_encode_protein", "hb_query_prepare_masked", "hb_query_set_state_mode",
    "hb_query_set_reset_interval", "hb_scan_noalloc", "hb_check_memory", "hb_cosched_run",
    "hb_cosearch",
    "hb_query_state_mode", "hb_scan_state_mode", "hb_batch_add_masked", "hb_scan_start", "hb_scan_wait", "hb_scan_results", "hb_scan_results_compact",
    "hb_scan_elapsed_ms", "hb_scan_free", "hb_search", "hb_diag_copy",
    "hb_gap_params_default", "hb_scan_gapped", "hb_scan_gapped_info", "hb_gapped_search",
    "hb_fasta_parse", "hb_hdb_gpu_seqs", "hb_hdb_layout", "hb_hdb_pack", "hb_hdbi_format",
    "hb_hdbi_parse", "hb_db_sort_volumes", "hb_hdb_write", "hb_db_upload_hdb", "hb_db_hdb_info", "hb_db_fetch",
    "hb_runtime_options_default", "hb_runtime_options_parse", "hb_hdb_split",
    "hb_balance_next_ratio", "hb_seg_params_default", "hb_seg_mask",
    "hb_batch_opts_default", "hb_batch_create", "hb_batch_add", "hb_batch_run",
    "hb_batch_results", "hb_batch_info_get", "hb_batch_free",
    "hb_serve_opts_default", "hb_serve_start", "hb_serve_submit", "hb_serve_submit_masked",
    "hb_serve_wait",
    "hb_serve_results", "hb_serve_stats_get", "hb_serve_stop",
]


class HBError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = _lib().hb_strerror(code).decode() if _LIB is not None else ERRORS.get(code, code)
        super().__init__(f"{what}: {ERRORS.get(code, code)} ({msg})")


class _BatchOpts(C.Structure):
    _fields_ = [("group_residues", C.c_int32), ("use_graph", C.c_int32),
                ("tune_tiles", C.c_uint32), ("waves_per_workgroup", C.c_int32),
                ("ring_bytes", C.c_int32), ("items_per_workgroup", C.c_int32)]


class _BatchInfo(C.Structure):
    _fields_ = [("n_queries", C.c_int32), ("n_groups", C.c_int32), ("n_classes", C.c_int32),
                ("group_residues", C.c_int32), ("kernel_ms", C.c_float), ("build_ms", C.c_float),
                ("lookup_hits", C.c_uint64), ("extensions", C.c_uint64), ("hsps", C.c_uint64),
                ("waves_per_workgroup", C.c_int32), ("ring_bytes", C.c_int32),
                ("items_per_workgroup", C.c_int32), ("tune_candidates", C.c_int32)]


class _ServeOpts(C.Structure):
    _fields_ = [("max_batch_queries", C.c_int32), ("max_inflight", C.c_int32),
                ("max_batch_residues", C.c_int64), ("mem_budget", C.c_int64),
                ("group_residues", C.c_int32), ("max_pending", C.c_int32),
                ("hsp_per_cell", C.c_double), ("linger_us", C.c_int64)]


class _ServeStats(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("batches", C.c_uint64), ("query_residues", C.c_uint64),
                ("peak_bytes", C.c_uint64), ("peak_inflight", C.c_int32),
                ("reruns", C.c_int32), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("gpu_ms", C.c_double),
                ("wall_ms", C.c_double)]


class _HdbVolume(C.Structure):
    _fields_ = [("name", C.c_void_p), ("name_len", C.c_uint32), ("bytes", C.c_uint64),
                ("volume_seqs", C.c_int32), ("gpu_seqs", C.c_int32),
                ("store_limit", C.c_int32), ("break_limit", C.c_int32),
                ("n_groups", C.c_int32), ("group_len", C.c_void_p),
                ("cpu_rest_bytes", C.c_uint64)]


class _HdbIndex(C.Structure):
    _fields_ = [("blocksx", C.c_int32), ("threadsx", C.c_int32), ("percentage", C.c_int32),
                ("n_volumes", C.c_int32)]


class _RuntimeOptions(C.Structure):
    _fields_ = [("gpu_ratio", C.c_float), ("fixed_ratio", C.c_int32),
                ("capability_rate", C.c_double), ("max_hits_per_subject", C.c_int32),
                ("mem_ecc", C.c_int32), ("load_balancer", C.c_int32),
                ("load_factor", C.c_int32), ("subject_chunk", C.c_int32),
                ("base_value", C.c_int32)]


class _HdbShare(C.Structure):
    _fields_ = [("gpu_seqs", C.c_int32), ("n_groups", C.c_int32), ("bytes", C.c_uint64)]


class _BalanceIn(C.Structure):
    _fields_ = [("volume", C.c_uint32), ("ratio", C.c_float), ("cpu_threads", C.c_uint32),
                ("gpu_cards", C.c_uint32), ("volume_bytes", C.c_uint64 * 3),
                ("ratio_ring", C.c_float * 2), ("hsp_ring", C.c_int64 * 2),
                ("gapped", C.c_int32), ("query_length", C.c_int32),
                ("avg_query_length", C.c_uint32), ("gpu_seconds", C.c_double),
                ("cpu_seconds", C.c_double)]


class _SegParams(C.Structure):
    _fields_ = [("window", C.c_int32), ("locut", C.c_double), ("hicut", C.c_double),
                ("maxbogus", C.c_int32), ("maxtrim", C.c_int32)]


class _Params(C.Structure):
    _fields_ = [("word_size", C.c_int32), ("threshold", C.c_int32), ("window", C.c_int32),
                ("matrix", C.c_void_p)]


_LIB = None
_SYNTH = None


def lib_path(name="libhbseed.so"):
    return os.path.join(_LIBDIR, name)


def _lib():
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"native library {path} is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        vp, i32, u32, u64 = C.c_void_p, C.c_int32, C.c_uint32, C.c_uint64
        pp = C.POINTER(C.c_void_p)
        L.hb_strerror.restype = C.c_char_p
        L.hb_strerror.argtypes = [C.c_int]
        L.hb_blosum62.restype = vp
        L.hb_create.argtypes = [C.c_int, pp]
        L.hb_destroy.argtypes = [vp]
        L.hb_mem_info.argtypes = [vp, C.POINTER(u64), C.POINTER(u64)]
        L.hb_synchronize.argtypes = [vp]
        L.hb_host_alloc.argtypes = [vp, u64, pp]
        L.hb_host_free.argtypes = [vp, vp]
        L.hb_db_upload.argtypes = [vp, vp, vp, u32, pp]
        L.hb_db_free.argtypes = [vp]
        L.hb_db_subjects.restype = u32
        L.hb_db_subjects.argtypes = [vp]
        L.hb_db_residues.restype = u64
        L.hb_db_residues.argtypes = [vp]
        L.hb_query_prepare.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, vp, pp]
        L.hb_query_prepare_masked.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, pp]
        L.hb_query_free.argtypes = [vp]
        L.hb_query_set_state_mode.argtypes = [vp, C.c_int]
        L.hb_query_set_reset_interval.argtypes = [vp, i32]
        L.hb_scan_noalloc.argtypes = [vp, vp, vp, vp, u32, vp, vp, vp]
        L.hb_check_memory.argtypes = [vp, vp, u64, u64, u32, u64, C.c_int64, C.POINTER(u64),
                                      C.POINTER(i32)]
        L.hb_cosched_run.argtypes = [C.c_char_p, vp, i32, i32, i32, u32, vp, vp, vp, vp, i32,
                                     C.POINTER(i32)]
        L.hb_cosearch.argtypes = [vp, C.c_char_p, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, i32,
                                  i32, C.c_int64, vp, vp, vp, vp, i32, C.POINTER(i32)]
        L.hb_query_state_mode.argtypes = [vp]
        L.hb_scan_state_mode.argtypes = [vp]
        L.hb_query_lut_info.argtypes = [vp, C.POINTER(u32), C.POINTER(u32)]
        L.hb_query_lut_fetch.argtypes = [vp, vp, vp, u32]
        L.hb_blastx_concat_len.restype = i32
        L.hb_blastx_concat_len.argtypes = [i32]
        L.hb_blastx_translate.argtypes = [C.c_char_p, i32, vp, vp, vp]
        L.hb_encode_protein.argtypes = [C.c_char_p, i32, vp]
        L.hb_scan_start.argtypes = [vp, vp, vp, pp]
        L.hb_scan_wait.argtypes = [vp]
        L.hb_scan_results.argtypes = [vp, vp, vp, u64, C.POINTER(u64)]
        L.hb_scan_results_compact.argtypes = [vp, vp, u64, vp, u64, vp]
        L.hb_scan_elapsed_ms.argtypes = [vp, C.POINTER(C.c_float)]
        L.hb_scan_free.argtypes = [vp]
        L.hb_diag_copy.argtypes = [vp, u64, C.c_int, C.POINTER(C.c_float)]
        L.hb_gap_params_default.argtypes = [vp]
        L.hb_gap_params_default.restype = None
        L.hb_scan_gapped.argtypes = [vp, vp, vp, vp, vp, u64, C.POINTER(u64)]
        L.hb_scan_gapped_info.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(u32)]
        L.hb_gapped_search.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, vp, u64, vp, vp, vp, vp,
                                       u64, C.POINTER(u64)]
        L.hb_fasta_parse.argtypes = [C.c_char_p, u64, vp, u64, vp, vp, vp, u32,
                                     C.POINTER(u64), C.POINTER(u32)]
        L.hb_hdb_gpu_seqs.restype = i32
        L.hb_hdb_gpu_seqs.argtypes = [i32]
        L.hb_hdb_layout.argtypes = [vp, i32, i32, i32, vp, i32, C.POINTER(i32), C.POINTER(u64)]
        L.hb_hdb_pack.argtypes = [vp, vp, i32, i32, i32, vp, i32, vp, u64]
        L.hb_db_sort_volumes.argtypes = [vp, i32, u64, u64, vp, vp, i32, C.POINTER(i32)]
        L.hb_hdb_write.argtypes = [C.c_char_p, vp, vp, i32, vp, vp, i32, i32, i32]
        L.hb_hdbi_format.argtypes = [vp, vp, vp, u64, C.POINTER(u64)]
        L.hb_hdbi_parse.argtypes = [C.c_char_p, u64, vp, vp, i32, vp, u64, C.POINTER(u64),
                                    C.POINTER(u64)]
        L.hb_db_upload_hdb.argtypes = [vp, vp, u64, i32, i32, i32, vp, i32, pp]
        L.hb_db_hdb_info.argtypes = [vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.hb_db_fetch.argtypes = [vp, vp, u64, vp]
        L.hb_runtime_options_default.argtypes = [vp]
        L.hb_runtime_options_default.restype = None
        L.hb_runtime_options_parse.argtypes = [C.c_char_p, u64, vp]
        L.hb_hdb_split.argtypes = [vp, i32, i32, u64, i32, i32, i32, C.POINTER(C.c_float),
                                   C.POINTER(i32), vp]
        L.hb_balance_next_ratio.argtypes = [vp, vp, C.POINTER(C.c_float)]
        L.hb_seg_params_default.argtypes = [vp]
        L.hb_seg_params_default.restype = None
        L.hb_seg_mask.argtypes = [vp, i32, i32, vp, vp, vp, vp]
        L.hb_batch_opts_default.argtypes = [vp]
        L.hb_batch_create.argtypes = [vp, vp, vp, vp, pp]
        L.hb_batch_add.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, C.POINTER(i32)]
        L.hb_batch_add_masked.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, vp, C.POINTER(i32)]
        L.hb_batch_run.argtypes = [vp]
        L.hb_batch_results.argtypes = [vp, i32, vp, u64, C.POINTER(u64)]
        L.hb_batch_info_get.argtypes = [vp, vp]
        L.hb_batch_free.argtypes = [vp]
        L.hb_serve_opts_default.argtypes = [vp]
        L.hb_serve_opts_default.restype = None
        L.hb_serve_start.argtypes = [vp, vp, vp, vp, pp]
        L.hb_serve_submit.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, C.POINTER(u64)]
        L.hb_serve_submit_masked.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, vp,
                                             C.POINTER(u64)]
        L.hb_serve_wait.argtypes = [vp, u64, i32]
        L.hb_serve_results.argtypes = [vp, u64, vp, u64, C.POINTER(u64)]
        L.hb_serve_stats_get.argtypes = [vp, vp]
        L.hb_serve_stop.argtypes = [vp]
        L.hb_search.argtypes = [vp, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, u64,
                                C.POINTER(u64)]
        _LIB = L
    return _LIB


def _synth():
    global _SYNTH
    if _SYNTH is None:
        path = lib_path("libhbsynth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"native library {path} is not built")
        S = C.CDLL(path)
        S.hbs_lognormal_lengths.restype = C.c_uint64
        S.hbs_lognormal_lengths.argtypes = [C.c_uint64, C.c_uint32, C.c_double, C.c_double,
                                            C.c_uint32, C.c_uint32, C.c_void_p]
        S.hbs_random_protein.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p]
        S.hbs_db_fill.argtypes = [C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_int32, C.c_double, C.c_int]
        S.hbs_homolog.restype = C.c_int32
        S.hbs_homolog.argtypes = [C.c_uint64, C.c_void_p, C.c_int32, C.c_double, C.c_int,
                                  C.c_void_p, C.c_int32]
        S.hbs_blastx_query.argtypes = [C.c_uint64, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p,
                                       C.c_int32, C.c_void_p]
        _SYNTH = S
    return _SYNTH


def _check(rc, what):
    if rc != HB_OK:
        raise HBError(rc, what)


def _ptr(a):
    return C.c_void_p(a.ctypes.data) if a is not None and a.size else None


def blosum62():
    p = _lib().hb_blosum62()
    return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_int8)), shape=(28, 28)).copy()


def encode_protein(text: str) -> np.ndarray:
    out = np.zeros(len(text), np.uint8)
    _check(_lib().hb_encode_protein(text.encode(), len(text), _ptr(out)), "encode_protein")
    return out


def parse_fasta(text):
    """Protein FASTA (str or bytes) -> (residues uint8, offsets uint64[n+1],
    titles list of str), through hb_fasta_parse."""
    if isinstance(text, str):
        text = text.encode()
    nr, ns = C.c_uint64(), C.c_uint32()
    rc = _lib().hb_fasta_parse(text, len(text), None, 0, None, None, None, 0, C.byref(nr),
                               C.byref(ns))
    if rc:
        raise HBError(rc, f"hb_fasta_parse (text offset {nr.value})")
    res = np.zeros(max(nr.value, 1), np.uint8)
    offs = np.zeros(ns.value + 1, np.uint64)
    tpos = np.zeros(max(ns.value, 1), np.uint64)
    tlen = np.zeros(max(ns.value, 1), np.uint32)
    _check(_lib().hb_fasta_parse(text, len(text), _ptr(res), len(res), _ptr(offs), _ptr(tpos),
                                 _ptr(tlen), ns.value, C.byref(nr), C.byref(ns)),
           "hb_fasta_parse")
    titles = [text[int(p):int(p) + int(n)].decode(errors="replace")
              for p, n in zip(tpos[:ns.value], tlen[:ns.value])]
    return res[:nr.value].copy(), offs, titles


def decode_protein(codes) -> str:
    return "".join(STDAA[int(c)] for c in codes)


# ------------------------------------------------------------------ queries
class QuerySpec:
    """A concatenated query: residues plus per-context offsets/lengths and
    per-context X-drop/cutoff (what NCBI's BLAST_SequenceBlk, BlastQueryInfo
    and BlastInitialWordParameters carry to H_BLAST_pre)."""

    def __init__(self, seq, ctx_off, ctx_len, xdrop=None, cutoff=None, lut_mask=None):
        self.seq = np.ascontiguousarray(seq, dtype=np.uint8)
        # residues that seed no lookup-table word (NCBI's masked segments)
        self.lut_mask = None if lut_mask is None else np.ascontiguousarray(lut_mask, np.uint8)
        self.ctx_off = np.ascontiguousarray(ctx_off, dtype=np.int32)
        self.ctx_len = np.ascontiguousarray(ctx_len, dtype=np.int32)
        n = len(self.ctx_off)
        self.xdrop = np.full(n, DEFAULT_XDROP, np.int32) if xdrop is None else \
            np.ascontiguousarray(xdrop, dtype=np.int32)
        self.cutoff = np.full(n, DEFAULT_CUTOFF, np.int32) if cutoff is None else \
            np.ascontiguousarray(cutoff, dtype=np.int32)

    @property
    def n_ctx(self):
        return len(self.ctx_off)


def blastp_query(seq, xdrop=DEFAULT_XDROP, cutoff=DEFAULT_CUTOFF) -> QuerySpec:
    seq = np.ascontiguousarray(seq, dtype=np.uint8)
    return QuerySpec(seq, [0], [len(seq)], [xdrop], [cutoff])


def blastx_query(nt: str, xdrop=DEFAULT_XDROP, cutoff=DEFAULT_CUTOFF) -> QuerySpec:
    L = _lib()
    n = L.hb_blastx_concat_len(len(nt))
    seq = np.zeros(n, np.uint8)
    off = np.zeros(6, np.int32)
    ln = np.zeros(6, np.int32)
    _check(L.hb_blastx_translate(nt.encode(), len(nt), _ptr(seq), _ptr(off), _ptr(ln)),
           "blastx_translate")
    return QuerySpec(seq, off, ln, [xdrop] * 6, [cutoff] * 6)


SEG_MASK_LETTER = 21   # X


def seg_mask(spec: QuerySpec, window=12, locut=2.2, hicut=2.5, maxbogus=2, maxtrim=100):
    """SEG low-complexity mask of a query's contexts (hb_seg_mask): uint8
    array, 1 inside a segment.  The defaults are NCBI's -seg 12 2.2 2.5."""
    p = _SegParams(window, locut, hicut, maxbogus, maxtrim)
    out = np.zeros(spec.seq.size, np.uint8)
    _check(_lib().hb_seg_mask(_ptr(spec.seq), spec.seq.size, spec.n_ctx, _ptr(spec.ctx_off),
                              _ptr(spec.ctx_len), C.byref(p), _ptr(out)), "hb_seg_mask")
    return out


def seg_query(spec: QuerySpec, hard=True, **seg_kw) -> QuerySpec:
    """The query with SEG applied as NCBI applies a filter: no lookup-table
    word inside a segment, and with hard masking (BLAST+'s default) the
    segment's residues read as X in extensions too."""
    m = seg_mask(spec, **seg_kw)
    seq = spec.seq.copy()
    if hard:
        seq[m != 0] = SEG_MASK_LETTER
    lut = m if spec.lut_mask is None else (m | (spec.lut_mask != 0)).astype(np.uint8)
    return QuerySpec(seq, spec.ctx_off, spec.ctx_len, spec.xdrop, spec.cutoff, lut_mask=lut)


# ------------------------------------------------------------------ runtime
class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        _check(_lib().hb_create(device, C.byref(h)), "hb_create")
        self.h = h

    def close(self):
        if self.h:
            _lib().hb_destroy(self.h)
            self.h = None

    def mem_info(self):
        f, t = C.c_uint64(), C.c_uint64()
        _check(_lib().hb_mem_info(self.h, C.byref(f), C.byref(t)), "hb_mem_info")
        return f.value, t.value

    def synchronize(self):
        _check(_lib().hb_synchronize(self.h), "hb_synchronize")

    def pinned(self, nbytes):
        """A page-locked host buffer (hb_host_alloc); .array is its uint8 view."""
        return PinnedBuffer(self, nbytes)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class PinnedBuffer:
    """Page-locked host memory from hb_host_alloc; free with close() (the
    numpy view must not be used afterwards)."""

    def __init__(self, ctx: Context, nbytes):
        self.ctx = ctx
        p = C.c_void_p()
        _check(_lib().hb_host_alloc(ctx.h, int(nbytes), C.byref(p)), "hb_host_alloc")
        self.p = p
        self.array = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)),
                                           shape=(max(int(nbytes), 1),))[:int(nbytes)]

    def close(self):
        if self.p:
            self.array = None
            _lib().hb_host_free(self.ctx.h, self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Database:
    """A resident subject database: packed arrays (hb_db_upload) or an H-BLAST
    .hDb volume unpacked on the device (Database.from_hdb)."""

    def __init__(self, ctx: Context, residues: np.ndarray = None, offsets: np.ndarray = None,
                 _handle=None):
        self.ctx = ctx
        self.oid_base = 0
        if _handle is not None:
            self.h = _handle
            self.n = _lib().hb_db_subjects(_handle)
            return
        residues = np.ascontiguousarray(residues, dtype=np.uint8)
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        self.n = len(offsets) - 1
        h = C.c_void_p()
        _check(_lib().hb_db_upload(ctx.h, _ptr(residues), _ptr(offsets), self.n, C.byref(h)),
               "hb_db_upload")
        self.h = h

    @classmethod
    def from_hdb(cls, ctx: Context, hdb, blocksx, threadsx, n_seqs, group_len):
        """One .hDb volume (bytes or uint8 array) -> resident database of its
        n_seqs sequences, unpacked on the device (hb_db_upload_hdb)."""
        buf = np.frombuffer(hdb, np.uint8) if isinstance(hdb, (bytes, bytearray)) else \
            np.ascontiguousarray(hdb, dtype=np.uint8)
        gl = np.ascontiguousarray(group_len, dtype=np.int32)
        h = C.c_void_p()
        _check(_lib().hb_db_upload_hdb(ctx.h, _ptr(buf), buf.size, blocksx, threadsx, n_seqs,
                                       _ptr(gl), gl.size, C.byref(h)), "hb_db_upload_hdb")
        return cls(ctx, _handle=h)

    def hdb_kernel_ms(self):
        """(lengths, copy) kernel times of the .hDb unpacking, in ms."""
        a, b = C.c_float(), C.c_float()
        _check(_lib().hb_db_hdb_info(self.h, C.byref(a), C.byref(b)), "hb_db_hdb_info")
        return a.value, b.value

    @property
    def residues(self):
        return _lib().hb_db_residues(self.h)

    def fetch(self):
        """(residues, offsets) of the resident database (hb_db_fetch)."""
        total = _lib().hb_db_residues(self.h)
        res = np.zeros(max(total, 1), np.uint8)
        offs = np.zeros(self.n + 1, np.uint64)
        _check(_lib().hb_db_fetch(self.h, _ptr(res), res.size, _ptr(offs)), "hb_db_fetch")
        return res[:total], offs

    def close(self):
        if self.h:
            _lib().hb_db_free(self.h)
            self.h = None


def hdb_gpu_seqs(volume_seqs):
    """Sequences of a volume that go to its .hDb (the rest stay on the CPU)."""
    return _lib().hb_hdb_gpu_seqs(int(volume_seqs))


def hdb_layout(offsets, n_seqs, blocksx=512, threadsx=192):
    """(group lengths int32, file size) of a .hDb holding sequences 0..n_seqs-1."""
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    ng, nb = C.c_int32(), C.c_uint64()
    _check(_lib().hb_hdb_layout(_ptr(offsets), n_seqs, blocksx, threadsx, None, 0,
                                C.byref(ng), C.byref(nb)), "hb_hdb_layout")
    gl = np.zeros(max(ng.value, 1), np.int32)
    _check(_lib().hb_hdb_layout(_ptr(offsets), n_seqs, blocksx, threadsx, _ptr(gl), gl.size,
                                C.byref(ng), C.byref(nb)), "hb_hdb_layout")
    return gl[:ng.value].copy(), nb.value


def hdb_pack(residues, offsets, n_seqs, blocksx=512, threadsx=192):
    """.hDb bytes (uint8 array) and group lengths of sequences 0..n_seqs-1."""
    residues = np.ascontiguousarray(residues, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    gl, nb = hdb_layout(offsets, n_seqs, blocksx, threadsx)
    out = np.zeros(max(nb, 1), np.uint8)
    _check(_lib().hb_hdb_pack(_ptr(residues), _ptr(offsets), n_seqs, blocksx, threadsx,
                              _ptr(gl), gl.size, _ptr(out), nb), "hb_hdb_pack")
    return out[:nb], gl


def hdbi_format(blocksx, threadsx, volumes, percentage=99):
    """.hDbI text for volumes (dicts with name, bytes, volume_seqs, gpu_seqs,
    store_limit, break_limit, group_len, cpu_rest_bytes)."""
    idx = _HdbIndex(blocksx, threadsx, percentage, len(volumes))
    arr = (_HdbVolume * max(len(volumes), 1))()
    keep = []
    for i, v in enumerate(volumes):
        name = v["name"].encode()
        gl = np.ascontiguousarray(v["group_len"], dtype=np.int32)
        keep += [name, gl]
        arr[i] = _HdbVolume(C.cast(C.c_char_p(name), C.c_void_p), len(name), v["bytes"],
                            v["volume_seqs"], v["gpu_seqs"], v["store_limit"], v["break_limit"],
                            gl.size, _ptr(gl), v["cpu_rest_bytes"])
    n = C.c_uint64()
    _check(_lib().hb_hdbi_format(C.byref(idx), arr, None, 0, C.byref(n)), "hb_hdbi_format")
    buf = C.create_string_buffer(n.value + 1)
    _check(_lib().hb_hdbi_format(C.byref(idx), arr, buf, n.value, C.byref(n)),
           "hb_hdbi_format")
    return buf.raw[:n.value].decode()


def hdbi_parse(text):
    """.hDbI text -> dict(blocksx, threadsx, percentage, volumes=[dict ...])."""
    if isinstance(text, str):
        text = text.encode()
    idx = _HdbIndex()
    ngl, eoff = C.c_uint64(), C.c_uint64()
    rc = _lib().hb_hdbi_parse(text, len(text), C.byref(idx), None, 0, None, 0, C.byref(ngl),
                              C.byref(eoff))
    if rc:
        raise HBError(rc, f"hb_hdbi_parse (text offset {eoff.value})")
    vols = (_HdbVolume * max(idx.n_volumes, 1))()
    gl = np.zeros(max(ngl.value, 1), np.int32)
    _check(_lib().hb_hdbi_parse(text, len(text), C.byref(idx), vols, idx.n_volumes, _ptr(gl),
                                gl.size, C.byref(ngl), C.byref(eoff)), "hb_hdbi_parse")
    base = C.cast(C.c_char_p(text), C.c_void_p).value
    out, g0 = [], 0
    for i in range(idx.n_volumes):
        v = vols[i]
        off = v.name - base
        out.append(dict(name=text[off:off + v.name_len].decode(), bytes=v.bytes,
                        volume_seqs=v.volume_seqs, gpu_seqs=v.gpu_seqs,
                        store_limit=v.store_limit, break_limit=v.break_limit,
                        group_len=gl[g0:g0 + v.n_groups].copy(),
                        cpu_rest_bytes=v.cpu_rest_bytes))
        g0 += v.n_groups
    return dict(blocksx=idx.blocksx, threadsx=idx.threadsx, percentage=idx.percentage,
                volumes=out)


def sort_volumes(offsets, max_file_bytes=1_000_000_000, max_letters=0):
    """Volume cut and length order of H-BLAST's makeblastdb -sort_volumes
    (hb_db_sort_volumes) -> (order, volume_seqs): the database's k-th
    sequence is input sequence order[k]."""
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(offsets) - 1
    order = np.zeros(max(n, 1), np.int32)
    nv = C.c_int32()
    _check(_lib().hb_db_sort_volumes(_ptr(offsets), n, max_file_bytes, max_letters, _ptr(order),
                                     None, 0, C.byref(nv)), "hb_db_sort_volumes")
    vs = np.zeros(max(nv.value, 1), np.int32)
    _check(_lib().hb_db_sort_volumes(_ptr(offsets), n, max_file_bytes, max_letters, _ptr(order),
                                     _ptr(vs), len(vs), C.byref(nv)), "hb_db_sort_volumes")
    return order[:n].copy(), vs[:nv.value].copy()


def write_hdb(prefix, residues, offsets, volume_seqs=None, blocksx=512, threadsx=192,
              order=None):
    """Writes an H-BLAST GPU database as CreateH_BLAST_database lays it out
    (hb_hdb_write): volume files <name>.hDb (one volume) or <name>.NN.hDb,
    and the index <prefix>.hDbI.  The database's k-th sequence is input
    sequence order[k] (default: the input order); sequences must ascend in
    length inside each volume; volume_seqs lists the volumes' sizes
    (default: one volume).  Returns the index dict (as hdbi_parse gives
    it)."""
    residues = np.ascontiguousarray(residues, dtype=np.uint8)
    offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(offsets) - 1
    vs = np.ascontiguousarray([n] if volume_seqs is None else list(volume_seqs), np.int32)
    if int(vs.sum()) != n:
        raise ValueError("volume sizes must add up to the number of sequences")
    ordp = None
    if order is not None:
        order = np.ascontiguousarray(order, np.int32)
        if len(order) != n:
            raise ValueError("order must list every sequence once")
        ordp = _ptr(order)
    _check(_lib().hb_hdb_write(os.fsencode(prefix), _ptr(residues), _ptr(offsets), n, ordp,
                               _ptr(vs), len(vs), blocksx, threadsx), "hb_hdb_write")
    with open(prefix + ".hDbI", "rb") as f:
        return hdbi_parse(f.read())


def build_hdb(prefix, residues, offsets, max_file_bytes=1_000_000_000, max_letters=0,
              blocksx=512, threadsx=192):
    """From packed sequences in input order (FASTA order) to an H-BLAST GPU
    database, as makeblastdb -sort_volumes followed by
    CreateH_BLAST_database builds one: volumes cut and length-sorted
    (sort_volumes), then written (write_hdb).  Returns (index dict, order)."""
    order, vs = sort_volumes(offsets, max_file_bytes, max_letters)
    idx = write_hdb(prefix, residues, offsets, vs, blocksx, threadsx, order=order)
    return idx, order


def open_hdb(ctx: Context, index_path):
    """Loads every volume an .hDbI lists (files beside the index) -> list of
    Database, each with oid_base = its first sequence's database oid."""
    with open(index_path, "rb") as f:
        idx = hdbi_parse(f.read())
    d = os.path.dirname(index_path)
    out = []
    for v in idx["volumes"]:
        with open(os.path.join(d, v["name"]), "rb") as f:
            data = f.read()
        db = Database.from_hdb(ctx, data, idx["blocksx"], idx["threadsx"], v["gpu_seqs"],
                               v["group_len"])
        db.oid_base = v["break_limit"] - v["volume_seqs"]
        out.append(db)
    return out


_OPTION_KEYS = [f[0] for f in _RuntimeOptions._fields_]


def runtime_options(text=None):
    """The H-BLAST_runtime_options settings as a dict: the values in force
    without a file (text None), or parsed from the file's text
    (Read_H_BLAST_runtime_options, ncbi_blast_files/functions.cpp:142-203)."""
    o = _RuntimeOptions()
    if text is None:
        _lib().hb_runtime_options_default(C.byref(o))
    else:
        if isinstance(text, str):
            text = text.encode()
        _check(_lib().hb_runtime_options_parse(text, len(text), C.byref(o)),
               "hb_runtime_options_parse")
    return {k: getattr(o, k) for k in _OPTION_KEYS}


def hdb_split(group_len, gpu_seqs, nbytes, blocksx, threadsx, cpu_threads, ratio=-1.0, fixed=1):
    """The GPU's share of a volume at a workload ratio (hb_hdb_split,
    get_GPU_DB_information) -> (dict(gpu_seqs, n_groups, bytes), ratio,
    fixed) with the ratio and flag as the split leaves them."""
    gl = np.ascontiguousarray(group_len, np.int32)
    r, f, out = C.c_float(ratio), C.c_int32(fixed), _HdbShare()
    _check(_lib().hb_hdb_split(_ptr(gl), gl.size, int(gpu_seqs), int(nbytes), blocksx, threadsx,
                               cpu_threads, C.byref(r), C.byref(f), C.byref(out)),
           "hb_hdb_split")
    return dict(gpu_seqs=out.gpu_seqs, n_groups=out.n_groups, bytes=out.bytes), r.value, f.value


def balance_next_ratio(options, volume, ratio, cpu_threads, volume_bytes, ratio_ring, hsp_ring,
                       gpu_seconds, cpu_seconds, gapped=False, query_length=0,
                       avg_query_length=0, gpu_cards=1):
    """The ratio for volume + 1 (hb_balance_next_ratio, get_new_gpudb_ratio)."""
    o = _RuntimeOptions(**options)
    b = _BalanceIn(volume, ratio, cpu_threads, gpu_cards, (C.c_uint64 * 3)(*volume_bytes),
                   (C.c_float * 2)(*ratio_ring), (C.c_int64 * 2)(*hsp_ring), int(bool(gapped)),
                   query_length, avg_query_length, gpu_seconds, cpu_seconds)
    out = C.c_float()
    _check(_lib().hb_balance_next_ratio(C.byref(o), C.byref(b), C.byref(out)),
           "hb_balance_next_ratio")
    return out.value


class _VolumeReport(C.Structure):
    _fields_ = [("oid_base", C.c_int64), ("volume_seqs", C.c_int32), ("gpu_seqs", C.c_int32),
                ("ratio", C.c_float), ("on_cpu", C.c_int32), ("need_bytes", C.c_uint64),
                ("gpu_seconds", C.c_double), ("cpu_seconds", C.c_double),
                ("n_hsp", C.c_uint64), ("launch_seq", C.c_uint64), ("cpu_seq", C.c_uint64)]


_CPU_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int64, C.c_int64)
_GPU_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_uint32,
                      C.c_void_p, C.c_uint64)
_FITS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32,
                       C.POINTER(C.c_uint64), C.POINTER(C.c_int32))
_LAUNCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                         C.c_int32, C.c_void_p)
_FETCH_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double),
                        C.POINTER(C.c_uint64))
_ABANDON_FN = C.CFUNCTYPE(None, C.c_void_p)


class _CoschedSteps(C.Structure):
    _fields_ = [("user", C.c_void_p), ("fits", _FITS_FN), ("launch", _LAUNCH_FN),
                ("fetch", _FETCH_FN), ("abandon", _ABANDON_FN)]


def _report_dicts(reports, n):
    return [{f: getattr(reports[i], f) for f, _ in _VolumeReport._fields_} for i in range(n)]


class _Trap:
    """Runs Python callbacks under C: the first exception is kept (the
    callback returns 1, which stops the loop) and raised afterwards."""

    def __init__(self):
        self.exc = None

    def call(self, fn, fail=1, always=False):
        def wrapped(*a):
            if self.exc is not None and not always:
                return fail
            try:
                r = fn(*a)
                return 0 if r is None else r
            except BaseException as e:   # noqa: BLE001 - re-raised after the C call
                self.exc = e
                return fail
        return wrapped

    def raise_if(self):
        if self.exc is not None:
            raise self.exc


def check_memory(ctx: Context, q, file_bytes, n_seqs, max_len, residues=0, budget=-1):
    """Device bytes a plain scan of q over a database (or .hDb share) needs
    and whether they fit (hb_check_memory): (need, fits)."""
    need, fits = C.c_uint64(), C.c_int32()
    _check(_lib().hb_check_memory(ctx.h, q.h, int(file_bytes), int(residues), int(n_seqs),
                                  int(max_len), int(budget), C.byref(need), C.byref(fits)),
           "hb_check_memory")
    return need.value, bool(fits.value)


def cosched_run(index_path, device, cpu_search, cpu_threads, options=None, gapped=False,
                query_length=0, avg_query_length=0):
    """The co-scheduled volume loop (hb_cosched_run) with Python device
    steps: device.fits(volume, share, max_len) -> (need, fits);
    device.launch(volume, data, share, blocksx, threadsx, group_len) -> 0 or
    an HB_ERR code; device.fetch(volume, oid_base) -> (gpu_seconds, n_hsp);
    device.abandon().  share = dict(gpu_seqs, n_groups, bytes).  Returns one
    report dict per volume."""
    trap = _Trap()

    def share_of(p):
        sh = C.cast(p, C.POINTER(_HdbShare)).contents
        return dict(gpu_seqs=sh.gpu_seqs, n_groups=sh.n_groups, bytes=sh.bytes)

    def fits(_u, v, sh, max_len, need, ok):
        n, f = device.fits(v, share_of(sh), max_len)
        need[0], ok[0] = int(n), int(bool(f))

    def launch(_u, v, hdb, sh, bx, tx, gl):
        share = share_of(sh)
        data = C.string_at(hdb, share["bytes"])
        glen = np.ctypeslib.as_array(C.cast(gl, C.POINTER(C.c_int32)),
                                     shape=(max(share["n_groups"], 1),))[:share["n_groups"]].copy()
        return device.launch(v, data, share, bx, tx, glen)

    def fetch(_u, v, oid_base, secs, nh):
        g, n = device.fetch(v, oid_base)
        secs[0], nh[0] = float(g), int(n)

    def abandon(_u):
        device.abandon()

    steps = _CoschedSteps(None, _FITS_FN(trap.call(fits)), _LAUNCH_FN(trap.call(launch)),
                          _FETCH_FN(trap.call(fetch)), _ABANDON_FN(trap.call(abandon, None, always=True)))
    cpu = _CPU_FN(trap.call(lambda _u, v, lo, hi: cpu_search(v, lo, hi)))
    opts = _RuntimeOptions(**options) if options is not None else None
    reports = (_VolumeReport * 1024)()
    n = C.c_int32()
    rc = _lib().hb_cosched_run(os.fsencode(index_path), None if opts is None else C.byref(opts),
                               cpu_threads, int(bool(gapped)), int(query_length),
                               int(avg_query_length), C.byref(steps), cpu, None, reports, 1024,
                               C.byref(n))
    trap.raise_if()
    _check(rc, "hb_cosched_run")
    return _report_dicts(reports, n.value)


def cosearch(ctx: Context, index_path, spec: QuerySpec, cpu_search, cpu_threads, options=None,
             gapped=False, device_budget=None, threshold=11, window=40, word_size=3,
             matrix=None):
    """Searches every volume of an .hDbI on the GPU and with the caller's CPU
    work side by side (hb_cosearch, the order of
    GPU_BLAST_PreliminarySearchEngine, ncbi_blast_files/functions.cpp:
    1409-1790): the device runs volume v + 1's share while cpu_search(v, lo,
    hi) searches database oids [lo, hi), the rest of volume v.  A share over
    device_budget bytes (default: the device's free memory), or whose
    upload runs out of device memory, goes to cpu_search whole.  Returns one
    dict per volume (hb_volume_report fields, plus stats and hsps of the
    device share with database oids)."""
    trap = _Trap()
    got = {}

    def gpu(_u, v, oid_base, st, ns, h, nh):
        stats = np.zeros(ns, STATS_DTYPE)
        hsps = np.zeros(nh, HSP_DTYPE)
        if ns:
            C.memmove(stats.ctypes.data, st, ns * STATS_DTYPE.itemsize)
        if nh:
            C.memmove(hsps.ctypes.data, h, nh * HSP_DTYPE.itemsize)
        got[v] = (stats, hsps)

    cpu = _CPU_FN(trap.call(lambda _u, v, lo, hi: cpu_search(v, lo, hi)))
    gfn = _GPU_FN(trap.call(gpu))
    p, keep = _params_struct(threshold, window, word_size, matrix)
    opts = _RuntimeOptions(**options) if options is not None else None
    reports = (_VolumeReport * 1024)()
    n = C.c_int32()
    mask = getattr(spec, "lut_mask", None)
    rc = _lib().hb_cosearch(ctx.h, os.fsencode(index_path), _ptr(spec.seq), len(spec.seq),
                            spec.n_ctx, _ptr(spec.ctx_off), _ptr(spec.ctx_len), _ptr(spec.xdrop),
                            _ptr(spec.cutoff), C.byref(p),
                            None if mask is None else _ptr(mask),
                            None if opts is None else C.byref(opts), cpu_threads,
                            int(bool(gapped)), -1 if device_budget is None else int(device_budget),
                            cpu, gfn, None, reports, 1024, C.byref(n))
    del keep
    trap.raise_if()
    _check(rc, "hb_cosearch")
    out = _report_dicts(reports, n.value)
    for v, rec in enumerate(out):
        rec["stats"], rec["hsps"] = got.get(v, (np.zeros(0, STATS_DTYPE), np.zeros(0, HSP_DTYPE)))
    return out


def _params_struct(threshold=11, window=40, word_size=3, matrix=None):
    p = _Params(word_size, threshold, window, None)
    keep = None
    if matrix is not None:
        keep = np.ascontiguousarray(matrix, dtype=np.int8).reshape(28 * 28)
        p.matrix = keep.ctypes.data
    return p, keep


class Query:
    def __init__(self, ctx: Context, spec: QuerySpec, threshold=11, window=40, word_size=3,
                 matrix=None):
        self.ctx = ctx
        self.spec = spec
        p, self._keep = _params_struct(threshold, window, word_size, matrix)
        h = C.c_void_p()
        _check(_lib().hb_query_prepare_masked(ctx.h, _ptr(spec.seq), len(spec.seq), spec.n_ctx,
                                              _ptr(spec.ctx_off), _ptr(spec.ctx_len),
                                              _ptr(spec.xdrop), _ptr(spec.cutoff), C.byref(p),
                                              _ptr(spec.lut_mask), C.byref(h)),
               "hb_query_prepare")
        self.h = h

    def set_state_mode(self, mode):
        """STATE_AUTO / STATE_LDS / STATE_GLOBAL: where scans keep the
        per-diagonal two-hit state (hb_query_set_state_mode; STATE_LDS16 is
        accepted as STATE_LDS)."""
        _check(_lib().hb_query_set_state_mode(self.h, mode), "hb_query_set_state_mode")

    def set_reset_interval(self, interval):
        """Testing aid: clear the diagonal state every `interval` accumulated
        subject residues per wavefront (hb_query_set_reset_interval; 0 = the
        placement's own limit)."""
        _check(_lib().hb_query_set_reset_interval(self.h, int(interval)),
               "hb_query_set_reset_interval")

    @property
    def state_mode(self):
        return _lib().hb_query_state_mode(self.h)

    def lut(self):
        ne, nu = C.c_uint32(), C.c_uint32()
        _check(_lib().hb_query_lut_info(self.h, C.byref(ne), C.byref(nu)), "lut_info")
        start = np.zeros(32769, np.uint32)
        ent = np.zeros(max(ne.value, 1), np.uint32)
        _check(_lib().hb_query_lut_fetch(self.h, _ptr(start), _ptr(ent), ne.value),
               "lut_fetch")
        return start, ent[:ne.value]

    def close(self):
        if self.h:
            _lib().hb_query_free(self.h)
            self.h = None


class ResultBuffers:
    """Host arrays reused by Scan.results / results_compact: each grows
    (to 1.25x the request) when a result does not fit, and is otherwise
    handed back as a view of its first n elements."""

    def __init__(self):
        self._a = {}

    def get(self, name, dtype, n):
        a = self._a.get(name)
        if a is None or a.dtype != dtype or len(a) < n:
            a = np.zeros(max(int(n * 1.25), 1), dtype)
            self._a[name] = a
        return a[:n]


class Scan:
    def __init__(self, ctx: Context, db: Database, q: Query):
        self.db = db
        h = C.c_void_p()
        _check(_lib().hb_scan_start(ctx.h, db.h, q.h, C.byref(h)), "hb_scan_start")
        self.h = h

    def wait(self):
        _check(_lib().hb_scan_wait(self.h), "hb_scan_wait")

    @property
    def state_mode(self):
        """Where this scan keeps its diagonal state (STATE_LDS / STATE_GLOBAL)."""
        return _lib().hb_scan_state_mode(self.h)

    def elapsed_ms(self):
        ms = C.c_float()
        _check(_lib().hb_scan_elapsed_ms(self.h, C.byref(ms)), "hb_scan_elapsed_ms")
        return ms.value

    def results(self, out=None):
        """-> (stats[n_subjects], hsps[n]) numpy structured arrays.  out: an
        optional ResultBuffers whose arrays are reused (grown when too small),
        the way the reference's caller keeps its result buffers across
        volumes; the returned arrays are then views of them."""
        n = C.c_uint64()
        _check(_lib().hb_scan_results(self.h, None, None, 0, C.byref(n)), "hb_scan_results")
        if out is None:
            stats = np.zeros(self.db.n, STATS_DTYPE)
            hsps = np.zeros(n.value, HSP_DTYPE)
        else:
            stats = out.get("stats", STATS_DTYPE, self.db.n)
            hsps = out.get("hsps", HSP_DTYPE, n.value)
        _check(_lib().hb_scan_results(self.h, _ptr(stats), _ptr(hsps), n.value, C.byref(n)),
               "hb_scan_results")
        return stats, hsps

    def results_compact(self, out=None):
        """-> (subjects[n_subjects with HSPs], hsps[n], totals dict): the
        compact results (hb_scan_results_compact).  out: as for results."""
        n = C.c_uint64()
        _check(_lib().hb_scan_results(self.h, None, None, 0, C.byref(n)), "hb_scan_results")
        cap = max(min(n.value, self.db.n), 1)   # subjects with HSPs <= both
        if out is None:
            subj = np.zeros(cap, SUBJ_HITS_DTYPE)
            hsps = np.zeros(n.value, HSP_DTYPE)
        else:
            subj = out.get("subj", SUBJ_HITS_DTYPE, cap)
            hsps = out.get("hsps", HSP_DTYPE, n.value)
        tot = (C.c_uint64 * 4)()
        _check(_lib().hb_scan_results_compact(self.h, _ptr(subj), len(subj), _ptr(hsps),
                                              n.value, tot), "hb_scan_results_compact")
        totals = dict(zip(("lookup_hits", "extensions", "n_hsp", "n_subjects"), list(tot)))
        return subj[:totals["n_subjects"]], hsps, totals

    def gapped(self, cutoffs, gap_open=DEFAULT_GAP_OPEN, gap_extend=DEFAULT_GAP_EXTEND,
               x_dropoff=DEFAULT_GAP_XDROP, out=None):
        """Preliminary gapped extension of this scan's HSPs (hb_scan_gapped)
        -> (stats[n_subjects], gapped hsps[n]) numpy structured arrays.  out:
        an optional ResultBuffers (as for results); the returned arrays are
        then views of its arrays."""
        gp = (C.c_int32 * 4)(gap_open, gap_extend, x_dropoff, 0)
        cut = np.ascontiguousarray(cutoffs, np.int32)
        n_ungapped = C.c_uint64()
        _check(_lib().hb_scan_results(self.h, None, None, 0, C.byref(n_ungapped)),
               "hb_scan_results")
        cap = max(n_ungapped.value, 1)
        if out is None:
            stats = np.zeros(self.db.n, GSTATS_DTYPE)
            ghsp = np.zeros(cap, GHSP_DTYPE)
        else:
            stats = out.get("gstats", GSTATS_DTYPE, self.db.n)
            ghsp = out.get("ghsps", GHSP_DTYPE, cap)
        n = C.c_uint64()
        _check(_lib().hb_scan_gapped(self.h, C.byref(gp), _ptr(cut), _ptr(stats), _ptr(ghsp),
                                     len(ghsp), C.byref(n)), "hb_scan_gapped")
        return stats, (ghsp[:n.value].copy() if out is None else ghsp[:n.value])

    def gapped_info(self):
        """-> (device ms of the last gapped stage, subjects run in the wide pass)"""
        ms, nw = C.c_float(), C.c_uint32()
        _check(_lib().hb_scan_gapped_info(self.h, C.byref(ms), C.byref(nw)),
               "hb_scan_gapped_info")
        return ms.value, nw.value

    def close(self):
        if self.h:
            _lib().hb_scan_free(self.h)
            self.h = None


class Batch:
    """Many queries in one batched scan (hb_batch_*)."""

    def __init__(self, ctx: Context, db: Database, group_residues=0, use_graph=True,
                 tune_tiles=0, threshold=11, window=40, word_size=3, matrix=None,
                 waves_per_workgroup=0, ring_bytes=0, items_per_workgroup=0):
        self.ctx, self.db = ctx, db
        p, self._keep = _params_struct(threshold, window, word_size, matrix)
        o = _BatchOpts(group_residues, 1 if use_graph else 0, tune_tiles, waves_per_workgroup,
                       ring_bytes, items_per_workgroup)
        h = C.c_void_p()
        _check(_lib().hb_batch_create(ctx.h, db.h, C.byref(p), C.byref(o), C.byref(h)),
               "hb_batch_create")
        self.h = h
        self.specs = []

    def add(self, spec: QuerySpec):
        qid = C.c_int32()
        _check(_lib().hb_batch_add_masked(self.h, _ptr(spec.seq), len(spec.seq), spec.n_ctx,
                                          _ptr(spec.ctx_off), _ptr(spec.ctx_len),
                                          _ptr(spec.xdrop), _ptr(spec.cutoff),
                                          _ptr(spec.lut_mask), C.byref(qid)), "hb_batch_add")
        self.specs.append(spec)
        return qid.value

    def run(self):
        _check(_lib().hb_batch_run(self.h), "hb_batch_run")

    def results(self, qid):
        n = C.c_uint64()
        _check(_lib().hb_batch_results(self.h, qid, None, 0, C.byref(n)), "hb_batch_results")
        out = np.zeros(n.value, HSP_DTYPE)
        _check(_lib().hb_batch_results(self.h, qid, _ptr(out), n.value, C.byref(n)),
               "hb_batch_results")
        return out

    def info(self):
        i = _BatchInfo()
        _check(_lib().hb_batch_info_get(self.h, C.byref(i)), "hb_batch_info_get")
        return {f: getattr(i, f) for f, _ in _BatchInfo._fields_}

    def close(self):
        if self.h:
            _lib().hb_batch_free(self.h)
            self.h = None


class Server:
    """The served form (hb_serve_*): submit queries from any thread, wait for
    each ticket, fetch its HSPs."""

    def __init__(self, ctx: Context, db: Database, max_batch_queries=None, max_inflight=None,
                 max_batch_residues=None, mem_budget=None, group_residues=None,
                 max_pending=None, hsp_per_cell=None, linger_us=None, threshold=11, window=40,
                 word_size=3, matrix=None):
        self.ctx, self.db = ctx, db
        p, self._keep = _params_struct(threshold, window, word_size, matrix)
        o = _ServeOpts()
        _lib().hb_serve_opts_default(C.byref(o))
        for k, v in (("max_batch_queries", max_batch_queries), ("max_inflight", max_inflight),
                     ("max_batch_residues", max_batch_residues), ("mem_budget", mem_budget),
                     ("group_residues", group_residues), ("max_pending", max_pending),
                     ("hsp_per_cell", hsp_per_cell), ("linger_us", linger_us)):
            if v is not None:
                setattr(o, k, v)
        h = C.c_void_p()
        _check(_lib().hb_serve_start(ctx.h, db.h, C.byref(p), C.byref(o), C.byref(h)),
               "hb_serve_start")
        self.h = h

    def submit(self, spec: QuerySpec):
        """Returns the request's ticket; raises HBError (code -8 busy, -4
        capacity) when the request is not admitted."""
        t = C.c_uint64()
        if spec.lut_mask is not None:
            _check(_lib().hb_serve_submit_masked(self.h, _ptr(spec.seq), len(spec.seq),
                                                 spec.n_ctx, _ptr(spec.ctx_off),
                                                 _ptr(spec.ctx_len), _ptr(spec.xdrop),
                                                 _ptr(spec.cutoff), _ptr(spec.lut_mask),
                                                 C.byref(t)), "hb_serve_submit_masked")
        else:
            _check(_lib().hb_serve_submit(self.h, _ptr(spec.seq), len(spec.seq), spec.n_ctx,
                                          _ptr(spec.ctx_off), _ptr(spec.ctx_len),
                                          _ptr(spec.xdrop), _ptr(spec.cutoff), C.byref(t)),
                   "hb_serve_submit")
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """True when ready, False on time-out."""
        rc = _lib().hb_serve_wait(self.h, ticket, timeout_ms)
        if rc == -8:
            return False
        _check(rc, "hb_serve_wait")
        return True

    def results(self, ticket):
        """HSPs of a finished request (releases the ticket)."""
        n = C.c_uint64()
        _check(_lib().hb_serve_results(self.h, ticket, None, 0, C.byref(n)), "hb_serve_results")
        out = np.zeros(n.value, HSP_DTYPE)
        _check(_lib().hb_serve_results(self.h, ticket, _ptr(out), n.value, C.byref(n)),
               "hb_serve_results")
        return out

    def stats(self):
        st = _ServeStats()
        _check(_lib().hb_serve_stats_get(self.h, C.byref(st)), "hb_serve_stats_get")
        return {f: getattr(st, f) for f, _ in _ServeStats._fields_}

    def stop(self):
        if self.h:
            _check(_lib().hb_serve_stop(self.h), "hb_serve_stop")
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


def search(ctx, db, spec: QuerySpec, **kw):
    q = Query(ctx, spec, **kw)
    try:
        s = Scan(ctx, db, q)
        try:
            return s.results()
        finally:
            s.close()
    finally:
        q.close()


def gapped_search(ctx, db, spec: QuerySpec, hsps, cutoffs, gap_open=DEFAULT_GAP_OPEN,
                  gap_extend=DEFAULT_GAP_EXTEND, x_dropoff=DEFAULT_GAP_XDROP, matrix=None):
    """Gapped stage on caller-held ungapped HSPs of one query (hb_gapped_search)
    -> (stats[n_subjects], gapped hsps[n])."""
    hsps = np.ascontiguousarray(hsps, HSP_DTYPE)
    cut = np.ascontiguousarray(cutoffs, np.int32)
    gp = (C.c_int32 * 4)(gap_open, gap_extend, x_dropoff, 0)
    mat = None if matrix is None else np.ascontiguousarray(matrix, np.int8)
    stats = np.zeros(db.n, GSTATS_DTYPE)
    out = np.zeros(max(len(hsps), 1), GHSP_DTYPE)
    n = C.c_uint64()
    _check(_lib().hb_gapped_search(ctx.h, db.h, _ptr(spec.seq), len(spec.seq), len(spec.ctx_off),
                                   _ptr(spec.ctx_off), _ptr(spec.ctx_len),
                                   None if mat is None else _ptr(mat), _ptr(hsps), len(hsps),
                                   C.byref(gp), _ptr(cut), _ptr(stats), _ptr(out), len(out),
                                   C.byref(n)), "hb_gapped_search")
    return stats, out[:n.value].copy()


# ------------------------------------------------------------------ synthetic data
def lognormal_lengths(seed, n, median, sigma, lo, hi):
    out = np.zeros(n, np.uint32)
    _synth().hbs_lognormal_lengths(seed, n, median, sigma, lo, hi, _ptr(out))
    return out


def random_protein(seed, n):
    out = np.zeros(n, np.uint8)
    _synth().hbs_random_protein(seed, n, _ptr(out))
    return out


def make_db(seed, n, median=300.0, sigma=0.6, lo=30, hi=5000, plant_query=None,
            plant_frac=0.0, threads=None):
    """Synthetic packed database -> (residues uint8, offsets uint64[n+1])."""
    lens = lognormal_lengths(seed, n, median, sigma, lo, hi)
    offs = np.zeros(n + 1, np.uint64)
    np.cumsum(lens, out=offs[1:])
    res = np.zeros(int(offs[-1]), np.uint8)
    pq = None if plant_query is None else np.ascontiguousarray(plant_query, np.uint8)
    threads = threads or min(16, os.cpu_count() or 1)
    _synth().hbs_db_fill(seed, n, _ptr(offs), _ptr(res), _ptr(pq),
                         0 if pq is None else len(pq), plant_frac, threads)
    return res, offs


def sort_by_length(residues, offsets):
    """Packed database -> (residues, offsets, order) with the sequences in
    ascending length order (stable), as an .hDb volume needs them."""
    offsets = np.asarray(offsets, np.uint64)
    lens = np.diff(offsets.astype(np.int64))
    order = np.argsort(lens, kind="stable")
    s_offs = np.zeros(len(order) + 1, np.uint64)
    np.cumsum(lens[order], out=s_offs[1:])
    starts = offsets[:-1].astype(np.int64)
    s_res = np.concatenate([residues[starts[i]:starts[i] + lens[i]] for i in order]) \
        if len(order) else np.zeros(0, np.uint8)
    return s_res.astype(np.uint8, copy=False), s_offs, order


def homolog(seed, src, identity, n_indel):
    src = np.ascontiguousarray(src, np.uint8)
    out = np.zeros(len(src) + 16, np.uint8)
    n = _synth().hbs_homolog(seed, _ptr(src), len(src), identity, n_indel, _ptr(out), len(out))
    return out[:n]


def blastx_nt(seed, nt_len, seg1, seg2):
    seg1 = np.ascontiguousarray(seg1, np.uint8)
    seg2 = np.ascontiguousarray(seg2, np.uint8)
    out = C.create_string_buffer(nt_len)
    rc = _synth().hbs_blastx_query(seed, nt_len, _ptr(seg1), len(seg1), _ptr(seg2), len(seg2),
                                   out)
    if rc:
        raise ValueError("coding regions do not fit the query")
    return out.raw.decode()
