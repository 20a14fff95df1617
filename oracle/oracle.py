"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker and by bench.py's cpu_baseline leg.  The product never imports it.

  textbook(packed, k)          closed-form unitigs (any 2 <= k <= 64)
  refpipe(packed, k, l, B)     serial restatement of the reference pipeline
                               (k = 15 mod 16), with per-phase times
  bucketize(...)               the reference's bucketing restated
  generate_permutation(l, s)   the reference's l-mer permutation restated
  lmer_order(code, l, seed)    the seeded bijection shared with the GPU
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "liboracle.so")

_U32P = C.POINTER(C.c_uint32)
_U64P = C.POINTER(C.c_uint64)


class OrcBuckets(C.Structure):
    _fields_ = [
        ("n_buckets", C.c_uint32),
        ("m", C.c_uint64),
        ("bucket_rows", _U64P),
        ("kmer_src", _U64P),
        ("cross_rows", _U64P),
        ("cross_lend", _U32P),
        ("cross_borrow", _U32P),
    ]


class OrcStats(C.Structure):
    _fields_ = [
        ("bucketing", C.c_double),
        ("assemble", C.c_double),
        ("coarsen", C.c_double),
        ("write_intra", C.c_double),
        ("minitigs", C.c_double),
        ("cross_finish", C.c_double),
        ("total", C.c_double),
        ("n_kmers", C.c_uint64),
        ("m_placements", C.c_uint64),
        ("n_cross", C.c_uint64),
        ("cross_offset", C.c_uint64),
        ("n_lines", C.c_uint64),
        ("buckets_wall", C.c_double),
        ("threads", C.c_uint64),
    ]


ORC_E = {-1: "ORC_E_ARG", -2: "ORC_E_NOMEM", -3: "ORC_E_FORMAT", -4: "ORC_E_SELFLOOP", -5: "ORC_E_DUP",
         -6: "ORC_E_HANG"}


class OracleError(RuntimeError):
    def __init__(self, rc):
        super().__init__(ORC_E.get(rc, str(rc)))
        self.rc = rc


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        L = C.CDLL(SO)
        L.orc_textbook.restype = C.c_int
        L.orc_textbook.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.POINTER(C.c_char_p), _U64P, _U64P, _U64P]
        L.orc_textbook_succ.restype = C.c_int
        L.orc_textbook_succ.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, _U64P]
        L.orc_refpipe.restype = C.c_int
        L.orc_refpipe.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                  C.POINTER(C.c_char_p), _U64P, C.POINTER(OrcStats)]
        L.orc_refpipe_par.restype = C.c_int
        L.orc_refpipe_par.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                      C.c_int, C.POINTER(C.c_char_p), _U64P, C.POINTER(OrcStats)]
        L.orc_bucketize.restype = C.c_int
        L.orc_bucketize.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                    C.POINTER(OrcBuckets)]
        L.orc_buckets_free.restype = None
        L.orc_buckets_free.argtypes = [C.POINTER(OrcBuckets)]
        L.orc_generate_permutation.restype = C.c_int
        L.orc_generate_permutation.argtypes = [C.c_uint32, C.c_uint64, C.c_void_p]
        L.orc_lmer_order.restype = C.c_uint32
        L.orc_lmer_order.argtypes = [C.c_uint32, C.c_uint32, C.c_uint64]
        L.orc_double_lmin.restype = None
        L.orc_double_lmin.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, _U32P, _U32P]
        L.orc_fnv_pref.restype = C.c_uint32
        L.orc_fnv_pref.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_fnv_suf.restype = C.c_uint32
        L.orc_fnv_suf.argtypes = [C.c_void_p, C.c_uint32]
        L.orc_parse_text.restype = C.c_int
        L.orc_parse_text.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.POINTER(_U32P), _U64P]
        L.orc_ref_load_kmers.restype = C.c_int
        L.orc_ref_load_kmers.argtypes = [C.c_char_p, C.c_size_t, C.c_uint32, C.POINTER(_U32P), _U64P]
        L.orc_pack.restype = C.c_int
        L.orc_pack.argtypes = [C.c_char_p, C.c_uint64, C.c_uint32, C.c_void_p]
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.c_void_p]
        _lib = L
    return _lib


def _arr(packed):
    return np.ascontiguousarray(packed, dtype=np.uint32)


def _take_text(ptr, n):
    L = lib()
    out = C.string_at(ptr, n) if n else b""
    L.orc_free(ptr)
    return out


def textbook(packed, k):
    """-> (text bytes, n_unitigs, n_cycles)"""
    L = lib()
    a = _arr(packed)
    t = C.c_char_p()
    tl, nu, nc = C.c_uint64(), C.c_uint64(), C.c_uint64()
    rc = L.orc_textbook(a.ctypes.data, a.shape[0], k, C.byref(t), C.byref(tl), C.byref(nu), C.byref(nc))
    if rc:
        raise OracleError(rc)
    return _take_text(t, tl.value), nu.value, nc.value


def textbook_succ(packed, k):
    L = lib()
    a = _arr(packed)
    out = np.zeros(max(a.shape[0], 1), np.uint64)
    rc = L.orc_textbook_succ(a.ctypes.data, a.shape[0], k, out.ctypes.data_as(_U64P))
    if rc:
        raise OracleError(rc)
    return out[: a.shape[0]]


def generate_permutation(l, seed):
    L = lib()
    out = np.zeros(4 ** l, np.uint32)
    rc = L.orc_generate_permutation(l, seed, out.ctypes.data)
    if rc:
        raise OracleError(rc)
    return out


def lmer_order(code, l, seed):
    return lib().orc_lmer_order(code, l, seed)


def lmer_order_table(l, seed):
    return np.array([lib().orc_lmer_order(c, l, seed) for c in range(4 ** l)], np.uint32)


def double_lmin(kmer_words, k, l, order):
    L = lib()
    a = _arr(kmer_words)
    o = _arr(order)
    p1, p2 = C.c_uint32(), C.c_uint32()
    L.orc_double_lmin(a.ctypes.data, k, l, o.ctypes.data, C.byref(p1), C.byref(p2))
    return p1.value, p2.value


def bucketize(packed, k, l, n_buckets, order):
    """-> dict with bucket_rows, kmer_src, cross_rows, cross_lend, cross_borrow"""
    L = lib()
    a = _arr(packed)
    o = _arr(order)
    b = OrcBuckets()
    rc = L.orc_bucketize(a.ctypes.data, a.shape[0], k, l, n_buckets, o.ctypes.data, C.byref(b))
    if rc:
        raise OracleError(rc)
    B = n_buckets
    ncr = b.cross_rows[B * B]
    res = {
        "bucket_rows": np.ctypeslib.as_array(b.bucket_rows, shape=(B + 1,)).copy(),
        "kmer_src": np.ctypeslib.as_array(b.kmer_src, shape=(max(b.m, 1),))[: b.m].copy(),
        "cross_rows": np.ctypeslib.as_array(b.cross_rows, shape=(B * B + 1,)).copy(),
        "cross_lend": np.ctypeslib.as_array(b.cross_lend, shape=(max(ncr, 1),))[:ncr].copy(),
        "cross_borrow": np.ctypeslib.as_array(b.cross_borrow, shape=(max(ncr, 1),))[:ncr].copy(),
    }
    L.orc_buckets_free(C.byref(b))
    return res


def refpipe(packed, k, l, n_buckets, order, threads=1):
    """-> (text bytes, stats dict); threads > 1 runs orc_refpipe_par (same bytes)"""
    L = lib()
    a = _arr(packed)
    o = _arr(order)
    t = C.c_char_p()
    tl = C.c_uint64()
    st = OrcStats()
    if threads > 1:
        rc = L.orc_refpipe_par(a.ctypes.data, a.shape[0], k, l, n_buckets, o.ctypes.data, threads, C.byref(t),
                               C.byref(tl), C.byref(st))
    else:
        rc = L.orc_refpipe(a.ctypes.data, a.shape[0], k, l, n_buckets, o.ctypes.data, C.byref(t), C.byref(tl),
                           C.byref(st))
    if rc:
        raise OracleError(rc)
    stats = {f: getattr(st, f) for f, _ in OrcStats._fields_}
    return _take_text(t, tl.value), stats


def fnv_pref(kmer_words):
    a = _arr(kmer_words)
    return lib().orc_fnv_pref(a.ctypes.data, a.shape[0])


def fnv_suf(kmer_words):
    a = _arr(kmer_words)
    return lib().orc_fnv_suf(a.ctypes.data, a.shape[0])


def parse_text(text, k):
    """The library's text rule (every line, whatever the count): uc_parse_kmer_text and uc_load_text."""
    return _parse(lib().orc_parse_text, text, k)


def ref_load_kmers(text, k):
    """The reference's chunked read loop (driver.cpp:257-354); OracleError
    with rc ORC_E_HANG where that loop would not end."""
    return _parse(lib().orc_ref_load_kmers, text, k)


def _parse(fn, text, k):
    L = lib()
    p = _U32P()
    n = C.c_uint64()
    rc = fn(text, len(text), k, C.byref(p), C.byref(n))
    if rc:
        raise OracleError(rc)
    w = (k + 15) // 16
    out = np.ctypeslib.as_array(p, shape=(max(n.value, 1) * w,))[: n.value * w].copy().reshape(n.value, w)
    L.orc_free(p)
    return out


def pack(chars: bytes, n, k):
    L = lib()
    w = (k + 15) // 16
    out = np.zeros((n, w), np.uint32)
    rc = L.orc_pack(chars, n, k, out.ctypes.data)
    if rc:
        raise OracleError(rc)
    return out
