"""MI355X-native unitig compaction (the compaction path of UnitigCoarsening).

The product is the C-ABI library ``_lib/libuc.so`` (declared in
``include/uc_api.h``) and the ``_lib/uc_main`` command-line drop-in.  This
module is a thin ctypes mirror of that ABI for Python callers and tests;
no compaction work happens in Python, and every call fails loudly when the
native library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "_lib")
LIB_PATH = os.path.join(LIB_DIR, "libuc.so")
SYNTH_PATH = os.path.join(LIB_DIR, "libuc_synth.so")
CLI_PATH = os.path.join(LIB_DIR, "uc_main")

UC_OK = 0
UC_E_ARG, UC_E_NOMEM, UC_E_HIP, UC_E_STATE, UC_E_IO, UC_E_FORMAT, UC_E_LIMIT = -1, -2, -3, -4, -5, -6, -7
UC_ALL_BUCKETS = 0xFFFFFFFF
UC_MODE_BATCHED = 0
UC_MODE_PLAIN = 1
UC_FLAG_PROFILE = 1
UC_FLAG_XCD_MAP = 2
UC_FLAG_KEEP_SRC = 4
UC_FLAG_DIRECT_SCATTER = 8
UC_FLAG_EAGER = 16
UC_FLAG_COUNT_PROBES = 32
UC_FLAG_LOW_MEMORY = 64
UC_SCRATCH_BYTES = 256 << 20
UC_TUNE_KEEP_BUCKETS = 1

F_LENT = 1
F_BORROW = 2


class UcError(RuntimeError):
    def __init__(self, status, msg=""):
        super().__init__(f"uc status {status}: {msg}")
        self.status = status


class UcOpts(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("seed", C.c_uint64),
        ("n_buckets", C.c_uint32),
        ("mode", C.c_uint32),
        ("load_pct", C.c_uint32),
        ("sample_log2", C.c_uint32),
        ("hbm_budget", C.c_uint64),
        ("flags", C.c_uint32),
        ("fine_slots", C.c_uint32),
        ("group_slots", C.c_uint32),
    ]


class UcBucketInfo(C.Structure):
    _fields_ = [
        ("n_kmers", C.c_uint64),
        ("m_placements", C.c_uint64),
        ("n_cross", C.c_uint64),
        ("n_buckets", C.c_uint32),
        ("n_sets", C.c_uint32),
        ("largest_bucket", C.c_uint64),
        ("n_fine_buckets", C.c_uint64),
        ("largest_fine_bucket", C.c_uint64),
        ("bucket_tile", C.c_uint32),
        ("fine_tile", C.c_uint32),
        ("lds_limit", C.c_uint32),
        ("n_oversized", C.c_uint32),
        ("oversized_slots", C.c_uint64),
        ("n_big_lds", C.c_uint64),
        ("n_slice_tasks", C.c_uint64),
        ("n_hbm_tables", C.c_uint64),
    ]


class UcCoarsenInfo(C.Structure):
    _fields_ = [
        ("n_unitigs", C.c_uint64),
        ("n_cycles", C.c_uint64),
        ("text_bytes", C.c_uint64),
        ("n_splitters", C.c_uint64),
        ("chain_rounds", C.c_uint32),
        ("n_runs", C.c_uint64),
    ]


class UcTiming(C.Structure):
    _fields_ = [
        ("bucketize_ms", C.c_float),
        ("edges_ms", C.c_float),
        ("coarsen_ms", C.c_float),
        ("emit_ms", C.c_float),
        ("total_ms", C.c_float),
        ("device_bytes", C.c_uint64),
        ("device_peak_bytes", C.c_uint64),
    ]


class UcTuneResult(C.Structure):
    _fields_ = [
        ("n_buckets", C.c_uint32),
        ("fine_slots", C.c_uint32),
        ("group_slots", C.c_uint32),
        ("sample_log2", C.c_uint32),
        ("load_pct", C.c_uint32),
        ("trials", C.c_uint32),
        ("default_ms", C.c_float),
        ("best_ms", C.c_float),
    ]


class UcServerOpts(C.Structure):
    _fields_ = [
        ("ctx_opts", UcOpts),
        ("n_slots", C.c_uint32),
        ("hbm_budget", C.c_uint64),
    ]


class UcPackedHeader(C.Structure):
    _fields_ = [
        ("magic", C.c_char * 8),
        ("k", C.c_uint32),
        ("words", C.c_uint32),
        ("n", C.c_uint64),
        ("reserved", C.c_uint64),
    ]


UC_JOB_QUEUED, UC_JOB_RUNNING, UC_JOB_DONE, UC_JOB_FAILED = 0, 1, 2, 3


class SynthParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64),
        ("genome_len", C.c_uint64),
        ("k", C.c_uint32),
        ("repeat_frac", C.c_double),
        ("seg_min", C.c_uint64),
        ("seg_max", C.c_uint64),
        ("repeat_sub", C.c_double),
        ("error_frac", C.c_double),
        ("copies_min", C.c_uint32),
        ("copies_max", C.c_uint32),
        ("div_min", C.c_double),
        ("div_max", C.c_double),
    ]


# every symbol of include/uc_api.h with (restype, argtypes)
_P = C.c_void_p
_U32P = C.POINTER(C.c_uint32)
_U64P = C.POINTER(C.c_uint64)
API = {
    "uc_default_opts": (None, [C.POINTER(UcOpts)]),
    "uc_ctx_create": (C.c_int, [C.POINTER(UcOpts), C.POINTER(_P)]),
    "uc_ctx_destroy": (C.c_int, [_P]),
    "uc_last_error": (C.c_char_p, [_P]),
    "uc_strerror": (C.c_char_p, [C.c_int]),
    "uc_compact": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int]),
    "uc_compact_loaded": (C.c_int, [_P, C.c_uint32]),
    "uc_compact_noalloc": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, _P, C.c_uint64, _U64P]),
    "uc_kernel_stat": (C.c_int, [_P, C.c_uint32, C.c_char_p, C.c_uint32, C.POINTER(C.c_double), _U64P]),
    "uc_kernel_stats_reset": (C.c_int, [_P]),
    "uc_kernel_stats_select": (C.c_int, [_P, C.c_char_p]),
    "uc_compact_sets": (C.c_int, [_P, _P, _U64P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int]),
    "uc_lmer_order_table": (C.c_int, [_P, C.c_uint32, _P, C.c_int]),
    "uc_load": (C.c_int, [_P, _P, _U64P, C.c_uint32, C.c_uint32, C.c_int]),
    "uc_bucketize": (C.c_int, [_P, C.c_uint32, _P, C.c_int]),
    "uc_bucket_info_get": (C.c_int, [_P, C.POINTER(UcBucketInfo)]),
    "uc_bucket_copy": (C.c_int, [_P, _U64P, _U32P, C.POINTER(C.c_uint8), _U32P, _U32P]),
    "uc_find_edges": (C.c_int, [_P, C.c_uint32]),
    "uc_succ_copy": (C.c_int, [_P, _U32P]),
    "uc_coarsen": (C.c_int, [_P]),
    "uc_coarsen_info_get": (C.c_int, [_P, C.POINTER(UcCoarsenInfo)]),
    "uc_emit": (C.c_int, [_P]),
    "uc_text_size": (C.c_uint64, [_P]),
    "uc_text_device_ptr": (C.c_void_p, [_P]),
    "uc_text_copy": (C.c_int, [_P, C.c_char_p, C.c_uint64]),
    "uc_set_text_offset": (C.c_uint64, [_P, C.c_uint32]),
    "uc_text_append_file": (C.c_int, [_P, C.c_char_p]),
    "uc_edge_probes": (C.c_int, [_P, _U64P]),
    "uc_parse_kmer_text": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(_U32P), _U64P]),
    "uc_read_kmer_file": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(_U32P), _U64P]),
    "uc_free": (None, [_P]),
    "uc_timing_get": (C.c_int, [_P, C.POINTER(UcTiming)]),
    "uc_sync": (C.c_int, [_P]),
    "uc_ctx_set_slice_probes": (C.c_int, [_P, C.c_uint32]),
    "uc_load_text": (C.c_int, [_P, C.c_char_p, C.c_uint64, C.c_uint32, C.c_int]),
    "uc_load_file": (C.c_int, [_P, C.c_char_p, C.c_uint32]),
    "uc_loaded_copy": (C.c_int, [_P, _U32P, C.c_uint64, _U64P]),
    "uc_loaded_device_ptr": (C.c_void_p, [_P]),
    "uc_write_packed_file": (C.c_int, [C.c_char_p, _U32P, C.c_uint64, C.c_uint32]),
    "uc_read_packed_header": (C.c_int, [C.c_char_p, C.POINTER(UcPackedHeader)]),
    "uc_load_packed_file": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_uint32)]),
    "uc_autotune": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(UcTuneResult)]),
    "uc_ctx_trim": (C.c_int, [_P]),
    "uc_ctx_held_bytes": (C.c_uint64, [_P]),
    "uc_job_bytes_estimate": (C.c_uint64, [C.c_uint64, C.c_uint32]),
    "uc_check_placements": (C.c_int, [C.c_uint64, C.c_uint64]),
    "uc_server_create": (C.c_int, [C.POINTER(UcServerOpts), C.POINTER(_P)]),
    "uc_server_destroy": (C.c_int, [_P]),
    "uc_submit": (C.c_int, [_P, _P, C.c_uint64, C.c_uint32, C.c_uint32, _U64P]),
    "uc_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int)]),
    "uc_wait": (C.c_int, [_P, C.c_uint64]),
    "uc_result": (C.c_int, [_P, C.c_uint64, C.c_char_p, C.c_uint64, _U64P, C.POINTER(C.c_double),
                            C.POINTER(C.c_double)]),
    "uc_release": (C.c_int, [_P, C.c_uint64]),
    "uc_server_stats": (C.c_int, [_P, _U64P, _U64P]),
}

_lib = None
_synth = None


def lib():
    """The compaction library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run __graft_entry__.build() or "
                               "python -m unitigcoarsening_amd.build")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in API.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def synth_lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"{SYNTH_PATH} missing: build first")
        S = C.CDLL(SYNTH_PATH)
        S.uc_synth_kmers.restype = C.c_int
        S.uc_synth_kmers.argtypes = [C.POINTER(SynthParams), C.POINTER(_U32P), _U64P]
        S.uc_synth_free.restype = None
        S.uc_synth_free.argtypes = [_P]
        S.uc_synth_format_text.restype = C.c_int
        S.uc_synth_format_text.argtypes = [_P, C.c_uint64, C.c_uint32, C.c_char_p, C.c_uint64, _U64P]
        _synth = S
    return _synth


def words_per_kmer(k: int) -> int:
    return (k + 15) // 16


def _check(rc, ctx=None):
    if rc != UC_OK:
        L = lib()
        msg = L.uc_strerror(rc).decode()
        if ctx:
            msg += ": " + (L.uc_last_error(ctx) or b"").decode()
        raise UcError(rc, msg)


# ---------------------------------------------------------------------------
# synthetic workloads and the text format


def synth_kmers(seed, genome_len, k, repeat_frac=0.0, seg_min=2000, seg_max=10000, repeat_sub=0.001,
                error_frac=0.0, copies=None, divergence=None) -> np.ndarray:
    """Distinct canonical k-mers of a seeded synthetic genome, reference packing
    (uint32 array of shape (n, ceil(k/16)))."""
    S = synth_lib()
    cmin, cmax = copies if copies else (0, 0)
    dmin, dmax = divergence if divergence else (0.0, 0.0)
    p = SynthParams(seed, genome_len, k, repeat_frac, seg_min, seg_max, repeat_sub, error_frac, cmin, cmax, dmin, dmax)
    ptr = _U32P()
    n = C.c_uint64()
    rc = S.uc_synth_kmers(C.byref(p), C.byref(ptr), C.byref(n))
    if rc != 0:
        raise RuntimeError(f"uc_synth_kmers failed ({rc})")
    w = words_per_kmer(k)
    out = np.ctypeslib.as_array(ptr, shape=(max(n.value, 1) * w,))[: n.value * w].copy().reshape(n.value, w)
    S.uc_synth_free(ptr)
    return out


def medium_model(seed, genome_len, k=31):
    """The repeat and error model of the medium workload: 5% repeats of
    2-10 kbp segments with 0.1% substitutions, 1% error k-mers."""
    return synth_kmers(seed, genome_len, k, repeat_frac=0.05, seg_min=2000, seg_max=10000, repeat_sub=0.001,
                       error_frac=0.01)


def large_model(seed, genome_len, k=63):
    """The repeat and error model of the large single instance: 45% repeat
    content in families of 300 bp - 6 kbp segments, copy number log-uniform
    2-1000, per-family divergence 0.5-15%; 0.5% error k-mers."""
    return synth_kmers(seed, genome_len, k, repeat_frac=0.45, seg_min=300, seg_max=6000, error_frac=0.005,
                       copies=(2, 1000), divergence=(0.005, 0.15))


def format_text(packed: np.ndarray, k: int) -> bytes:
    """Reference text format: count line, then one k-mer per line."""
    S = synth_lib()
    packed = np.ascontiguousarray(packed, dtype=np.uint32)
    n = packed.shape[0]
    need = C.c_uint64()
    S.uc_synth_format_text(packed.ctypes.data, n, k, None, 0, C.byref(need))
    buf = C.create_string_buffer(need.value)
    rc = S.uc_synth_format_text(packed.ctypes.data, n, k, buf, need.value, C.byref(need))
    if rc != 0:
        raise RuntimeError("format_text failed")
    return buf.raw[: need.value]


def parse_text(text: bytes, k: int) -> np.ndarray:
    """k-mer text -> reference packing via the library's host parser."""
    L = lib()
    ptr = _U32P()
    n = C.c_uint64()
    _check(L.uc_parse_kmer_text(text, len(text), k, C.byref(ptr), C.byref(n)))
    w = words_per_kmer(k)
    out = np.ctypeslib.as_array(ptr, shape=(max(n.value, 1) * w,))[: n.value * w].copy().reshape(n.value, w)
    L.uc_free(ptr)
    return out


def decode(packed: np.ndarray, k: int) -> list:
    """Reference packing -> list of k-mer strings (for tests and small data)."""
    packed = np.asarray(packed, dtype=np.uint32)
    out = []
    for row in packed:
        out.append("".join("ACGT"[(int(row[i // 16]) >> (2 * (i % 16))) & 3] for i in range(k)))
    return out


def encode(kmers, k: int) -> np.ndarray:
    """k-mer strings -> reference packing."""
    w = words_per_kmer(k)
    code = {"A": 0, "C": 1, "G": 2, "T": 3}
    out = np.zeros((len(kmers), w), dtype=np.uint32)
    for j, s in enumerate(kmers):
        for i in range(k):
            out[j, i // 16] |= code.get(s[i], 0) << (2 * (i % 16))
    return out


# ---------------------------------------------------------------------------
# the compactor


class Compactor:
    """One device context (uc_ctx).  Not thread-safe: one host thread per object."""

    def __init__(self, device=0, seed=0x5EED, n_buckets=64, mode=UC_MODE_BATCHED, load_pct=50, sample_log2=0,
                 hbm_budget=0, profile=False, xcd_map=False, keep_src=False, fine_slots=0, group_slots=0,
                 direct_scatter=False, eager=False, count_probes=False, low_memory=False):
        L = lib()
        o = UcOpts()
        L.uc_default_opts(C.byref(o))
        o.device, o.seed, o.n_buckets, o.mode = device, seed, n_buckets, mode
        o.load_pct, o.sample_log2, o.hbm_budget = load_pct, sample_log2, hbm_budget
        o.flags = ((UC_FLAG_PROFILE if profile else 0) | (UC_FLAG_XCD_MAP if xcd_map else 0)
                   | (UC_FLAG_KEEP_SRC if keep_src else 0) | (UC_FLAG_DIRECT_SCATTER if direct_scatter else 0)
                   | (UC_FLAG_EAGER if eager else 0) | (UC_FLAG_COUNT_PROBES if count_probes else 0)
                   | (UC_FLAG_LOW_MEMORY if low_memory else 0))
        o.fine_slots, o.group_slots = fine_slots, group_slots
        self.opts = o
        self._ctx = _P()
        _check(L.uc_ctx_create(C.byref(o), C.byref(self._ctx)))
        self._keep = None

    def close(self):
        if self._ctx:
            lib().uc_ctx_destroy(self._ctx)
            self._ctx = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- whole jobs ---------------------------------------------------------
    def compact(self, packed, k, l, device_ptr=None, n=None):
        """Run the whole path; returns the unitig text (bytes)."""
        L = lib()
        if device_ptr is not None:
            _check(L.uc_compact(self._ctx, C.c_void_p(device_ptr), n, k, l, 1), self._ctx)
        else:
            packed = np.ascontiguousarray(packed, dtype=np.uint32)
            self._keep = packed
            _check(L.uc_compact(self._ctx, packed.ctypes.data, packed.shape[0], k, l, 0), self._ctx)
        return self.text()

    def compact_sets(self, sets, k, l):
        """Batched form over independent k-mer sets; returns one text per set."""
        L = lib()
        arrs = [np.ascontiguousarray(s, dtype=np.uint32).reshape(-1, words_per_kmer(k)) for s in sets]
        offs = np.zeros(len(arrs) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum([a.shape[0] for a in arrs])
        allk = np.ascontiguousarray(np.concatenate(arrs, axis=0)) if arrs else np.zeros((0, 1), np.uint32)
        self._keep = allk
        _check(L.uc_compact_sets(self._ctx, allk.ctypes.data, offs.ctypes.data_as(_U64P), len(arrs), k, l, 0),
               self._ctx)
        t = self.text()
        bounds = [L.uc_set_text_offset(self._ctx, i) for i in range(len(arrs) + 1)]
        return [t[bounds[i]:bounds[i + 1]] for i in range(len(arrs))]

    # -- phases -------------------------------------------------------------
    def load(self, packed, k, set_sizes=None):
        L = lib()
        packed = np.ascontiguousarray(packed, dtype=np.uint32)
        self._keep = packed
        sizes = [packed.shape[0]] if set_sizes is None else list(set_sizes)
        offs = np.zeros(len(sizes) + 1, dtype=np.uint64)
        offs[1:] = np.cumsum(sizes)
        _check(L.uc_load(self._ctx, packed.ctypes.data, offs.ctypes.data_as(_U64P), len(sizes), k, 0), self._ctx)

    def load_text(self, text: bytes, k):
        """Parse the k-mer text on the device (uc_load_text) and load it."""
        _check(lib().uc_load_text(self._ctx, text, len(text), k, 0), self._ctx)

    def load_file(self, path, k):
        """Read and parse a k-mer file on the device (uc_load_file)."""
        _check(lib().uc_load_file(self._ctx, os.fsencode(path), k), self._ctx)

    def load_packed_file(self, path):
        """Stream a packed k-mer file into HBM (uc_load_packed_file); returns its k."""
        k = C.c_uint32()
        _check(lib().uc_load_packed_file(self._ctx, os.fsencode(path), C.byref(k)), self._ctx)
        return k.value

    def loaded(self, k):
        """The loaded k-mers in the reference packing, shape (n, ceil(k/16))."""
        L = lib()
        n = C.c_uint64()
        _check(L.uc_loaded_copy(self._ctx, None, 0, C.byref(n)), self._ctx)
        out = np.zeros((max(n.value, 1), words_per_kmer(k)), np.uint32)
        _check(L.uc_loaded_copy(self._ctx, out.ctypes.data_as(_U32P), n.value, None), self._ctx)
        return out[:n.value]

    def compact_loaded(self, l):
        """Whole job on the k-mers of the last load(), already in device memory."""
        _check(lib().uc_compact_loaded(self._ctx, l), self._ctx)

    def kernel_stats(self):
        """{kernel name: (total ms, launches)} since the last reset (profile=True)."""
        L = lib()
        out = {}
        i = 0
        name = C.create_string_buffer(64)
        ms = C.c_double()
        nl = C.c_uint64()
        while L.uc_kernel_stat(self._ctx, i, name, 64, C.byref(ms), C.byref(nl)) == UC_OK:
            out[name.value.decode()] = (ms.value, nl.value)
            i += 1
        return out

    def kernel_stats_reset(self):
        _check(lib().uc_kernel_stats_reset(self._ctx), self._ctx)

    def kernel_stats_select(self, name=None):
        """Time only the named kernel (None: every kernel) from now on."""
        _check(lib().uc_kernel_stats_select(self._ctx, (name or "").encode()), self._ctx)

    def bucketize(self, l, order_table=None):
        """order_table: None (inline seeded bijection) or a host array of 4^l orders."""
        if order_table is None:
            _check(lib().uc_bucketize(self._ctx, l, None, 0), self._ctx)
        else:
            t = np.ascontiguousarray(order_table, dtype=np.uint32)
            assert t.shape[0] == 4 ** l
            _check(lib().uc_bucketize(self._ctx, l, t.ctypes.data, 0), self._ctx)

    def lmer_order_table(self, l):
        out = np.zeros(4 ** l, np.uint32)
        _check(lib().uc_lmer_order_table(self._ctx, l, out.ctypes.data, 0), self._ctx)
        return out

    def bucket_info(self):
        i = UcBucketInfo()
        _check(lib().uc_bucket_info_get(self._ctx, C.byref(i)), self._ctx)
        return i

    def bucket_arrays(self):
        i = self.bucket_info()
        G = i.n_sets * i.n_buckets
        rows = np.zeros(G + 1, np.uint64)
        src = np.zeros(max(i.m_placements, 1), np.uint32)
        flags = np.zeros(max(i.m_placements, 1), np.uint8)
        ph = np.zeros(max(i.n_cross, 1), np.uint32)
        pb = np.zeros(max(i.n_cross, 1), np.uint32)
        _check(lib().uc_bucket_copy(self._ctx, rows.ctypes.data_as(_U64P), src.ctypes.data_as(_U32P),
                                    flags.ctypes.data_as(C.POINTER(C.c_uint8)), ph.ctypes.data_as(_U32P),
                                    pb.ctypes.data_as(_U32P)), self._ctx)
        m, x = i.m_placements, i.n_cross
        return rows, src[:m], flags[:m], ph[:x], pb[:x]

    def find_edges(self, bucket=UC_ALL_BUCKETS):
        _check(lib().uc_find_edges(self._ctx, bucket), self._ctx)

    def succ(self):
        i = self.bucket_info()
        out = np.zeros(max(i.m_placements, 1), np.uint32)
        _check(lib().uc_succ_copy(self._ctx, out.ctypes.data_as(_U32P)), self._ctx)
        return out[: i.m_placements]

    def coarsen(self):
        _check(lib().uc_coarsen(self._ctx), self._ctx)

    def coarsen_info(self):
        i = UcCoarsenInfo()
        _check(lib().uc_coarsen_info_get(self._ctx, C.byref(i)), self._ctx)
        return i

    def emit(self):
        _check(lib().uc_emit(self._ctx), self._ctx)

    # -- results ------------------------------------------------------------
    def text(self) -> bytes:
        L = lib()
        n = L.uc_text_size(self._ctx)
        buf = np.empty(max(n, 1), np.uint8)
        _check(L.uc_text_copy(self._ctx, buf.ctypes.data_as(C.c_char_p), n), self._ctx)
        return buf[:n].tobytes()

    def text_size(self):
        return lib().uc_text_size(self._ctx)

    def edge_probes(self):
        """Hash-table slot inspections of the last edge phase (uc_edge_probes)."""
        v = C.c_uint64()
        _check(lib().uc_edge_probes(self._ctx, C.byref(v)), self._ctx)
        return v.value

    def append_text_file(self, path):
        """Append the unitig text to the file at `path` (uc_text_append_file)."""
        _check(lib().uc_text_append_file(self._ctx, os.fsencode(path)), self._ctx)

    def timing(self):
        t = UcTiming()
        _check(lib().uc_timing_get(self._ctx, C.byref(t)), self._ctx)
        return t

    def sync(self):
        _check(lib().uc_sync(self._ctx), self._ctx)

    def set_slice_probes(self, probes):
        """Most probes of one slice-table insert (uc_ctx_set_slice_probes)."""
        _check(lib().uc_ctx_set_slice_probes(self._ctx, probes), self._ctx)

    def autotune(self, l, reps=2, keep_buckets=False):
        """Tune the settings on the loaded job (uc_autotune); returns the choice as a dict."""
        r = UcTuneResult()
        _check(lib().uc_autotune(self._ctx, l, reps, UC_TUNE_KEEP_BUCKETS if keep_buckets else 0, C.byref(r)),
               self._ctx)
        return {f: getattr(r, f) for f, _ in UcTuneResult._fields_}

    def trim(self):
        """Free every device buffer the context holds (uc_ctx_trim)."""
        _check(lib().uc_ctx_trim(self._ctx), self._ctx)

    def held_bytes(self):
        return lib().uc_ctx_held_bytes(self._ctx)


def write_packed_file(path, packed, k):
    """Packed binary k-mer file (uc_write_packed_file)."""
    packed = np.ascontiguousarray(packed, dtype=np.uint32).reshape(-1, words_per_kmer(k))
    _check(lib().uc_write_packed_file(os.fsencode(path), packed.ctypes.data_as(_U32P), packed.shape[0], k))


def read_packed_header(path):
    h = UcPackedHeader()
    _check(lib().uc_read_packed_header(os.fsencode(path), C.byref(h)))
    return {"k": h.k, "words": h.words, "n": h.n}


def job_bytes_estimate(n, k):
    """Upper bound of the device bytes one job of n k-mers needs (uc_job_bytes_estimate)."""
    return lib().uc_job_bytes_estimate(n, k)


class Server:
    """Request loop (uc_server): n_slots worker contexts, FIFO admission under an HBM budget.

    submit() copies the k-mers and returns a job id at once; result() waits and returns
    (text, queue_ms, run_ms).  Safe to call from several Python threads."""

    def __init__(self, n_slots=4, hbm_budget=0, device=0, seed=0x5EED, n_buckets=64, sample_log2=0):
        L = lib()
        o = UcServerOpts()
        L.uc_default_opts(C.byref(o.ctx_opts))
        o.ctx_opts.device, o.ctx_opts.seed = device, seed
        o.ctx_opts.n_buckets, o.ctx_opts.sample_log2 = n_buckets, sample_log2
        o.n_slots, o.hbm_budget = n_slots, hbm_budget
        self._srv = _P()
        _check(L.uc_server_create(C.byref(o), C.byref(self._srv)))

    def close(self):
        if self._srv:
            lib().uc_server_destroy(self._srv)
            self._srv = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, packed, k, l):
        packed = np.ascontiguousarray(packed, dtype=np.uint32)
        job = C.c_uint64()
        _check(lib().uc_submit(self._srv, packed.ctypes.data, packed.shape[0], k, l, C.byref(job)))
        return job.value

    def poll(self, job):
        st = C.c_int()
        _check(lib().uc_poll(self._srv, job, C.byref(st)))
        return st.value

    def result(self, job, release=True, text=True):
        """(text, queued ms, running ms) of a finished job; with text=False the
        text stays in the server (only its size is returned in its place)."""
        L = lib()
        L.uc_wait(self._srv, job)
        size, qms, rms = C.c_uint64(), C.c_double(), C.c_double()
        rc = L.uc_result(self._srv, job, None, 0, C.byref(size), C.byref(qms), C.byref(rms))
        out = size.value
        if rc == UC_OK and text:
            buf = np.empty(max(size.value, 1), np.uint8)
            rc = L.uc_result(self._srv, job, buf.ctypes.data_as(C.c_char_p), size.value, None, None, None)
            out = buf[:size.value].tobytes()
        if release:
            L.uc_release(self._srv, job)
        _check(rc)
        return out, qms.value, rms.value

    def stats(self):
        peak, done = C.c_uint64(), C.c_uint64()
        _check(lib().uc_server_stats(self._srv, C.byref(peak), C.byref(done)))
        return {"peak_held_bytes": peak.value, "jobs_done": done.value}


def sorted_lines(text: bytes) -> list:
    """Unitig text -> sorted list of lines (order-free comparison)."""
    return sorted(text.decode().splitlines())


_COMP = str.maketrans("ACGT", "TGCA")


def canonical_sorted(text: bytes) -> list:
    """Each unitig in canonical orientation (min of it and its reverse
    complement), then sorted: the multiset comparison of unitig sets."""
    out = []
    for s in text.decode().splitlines():
        r = s.translate(_COMP)[::-1]
        out.append(min(s, r))
    return sorted(out)
