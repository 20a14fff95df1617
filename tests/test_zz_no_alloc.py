"""uc_compact_noalloc (every buffer carved from the library's static device
scratch, the text written to the caller's buffer) gives the same bytes as
uc_compact on the same input, on two calls in a row, the second smaller;
both texts also equal the textbook oracle's unitig set."""
import ctypes as C

import numpy as np
import pytest

import oracle
import unitigcoarsening_amd as uc

pytestmark = pytest.mark.gpu


def _hip():
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


class DevBuf:
    """A device buffer the test allocates (the entry point must not)."""

    def __init__(self, hip, nbytes):
        self.hip, self.n = hip, max(int(nbytes), 1)
        self.p = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), self.n) == 0

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.n
        assert self.hip.hipMemcpy(self.p, arr.ctypes.data, arr.nbytes, 1) == 0  # host to device

    def get(self, nbytes):
        out = np.empty(nbytes, dtype=np.uint8)
        if nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, self.p, nbytes, 2) == 0  # device to host
        return out.tobytes()

    def free(self):
        if self.p:
            self.hip.hipFree(self.p)
            self.p = C.c_void_p()


def noalloc(c, hip, kmers, k, l, cap=None, fill=0x5A):
    """One uc_compact_noalloc call on a device copy of kmers; returns (status, size, text)."""
    L = uc.lib()
    kmers = np.ascontiguousarray(kmers, dtype=np.uint32)
    n = kmers.shape[0]
    src = DevBuf(hip, kmers.nbytes)
    src.put(kmers)
    size = C.c_uint64(0)
    if cap is None:
        cap = n * (k + 1) + 64  # n isolated k-mers: the most text a job can have
    out = DevBuf(hip, cap)
    try:
        # garbage in the output buffer: the call must write every text byte
        assert hip.hipMemset(out.p, fill, out.n) == 0
        rc = L.uc_compact_noalloc(c._ctx, src.p, n, k, l, out.p, cap, C.byref(size))
        text = out.get(size.value) if rc == uc.UC_OK else b""
        return rc, size.value, text
    finally:
        src.free()
        out.free()


def reference_text(kmers, k, l, n_buckets):
    with uc.Compactor(n_buckets=n_buckets) as c:
        return c.compact(kmers, k, l)


@pytest.mark.parametrize("k,l,sizes", [(31, 11, (600_000, 60_000)), (63, 13, (300_000, 40_000))])
def test_noalloc_matches_compact_twice(k, l, sizes):
    hip = _hip()
    big = uc.medium_model(11, sizes[0], k)
    small = uc.medium_model(12, sizes[1], k)
    assert big.shape[0] > small.shape[0] > 0
    with uc.Compactor(n_buckets=64) as c:
        for kmers in (big, small):
            want = reference_text(kmers, k, l, 64)
            rc, size, got = noalloc(c, hip, kmers, k, l)
            assert rc == uc.UC_OK, uc.lib().uc_last_error(c._ctx)
            assert size == len(want)
            assert got == want  # byte for byte: same slots, same line order
            text, _, _ = oracle.textbook(kmers, k)
            assert uc.sorted_lines(got) == uc.sorted_lines(text)
        # the context still runs ordinary jobs, on its own buffers
        assert c.compact(small, k, l) == reference_text(small, k, l, 64)


def test_noalloc_short_text_buffer_and_limits():
    hip = _hip()
    L = uc.lib()
    kmers = uc.medium_model(13, 50_000, 31)
    want = reference_text(kmers, 31, 11, 64)
    with uc.Compactor(n_buckets=64) as c:
        # a buffer one byte short: UC_E_NOMEM, with the size the text needs
        rc, size, _ = noalloc(c, hip, kmers, 31, 11, cap=len(want) - 1)
        assert rc == uc.UC_E_NOMEM
        assert size == len(want)
        # far beyond the scratch: refused before anything runs (the buffers
        # are real but small; the host check stops the call before any use)
        tiny = DevBuf(hip, 256)
        try:
            size = C.c_uint64(7)
            rc = L.uc_compact_noalloc(c._ctx, tiny.p, uc.UC_SCRATCH_BYTES, 31, 11, tiny.p, 256, C.byref(size))
            assert rc == uc.UC_E_LIMIT and size.value == 0
            # bad arguments
            assert L.uc_compact_noalloc(c._ctx, None, 5, 31, 11, tiny.p, 256, C.byref(size)) == uc.UC_E_ARG
            assert L.uc_compact_noalloc(c._ctx, tiny.p, 5, 1, 11, tiny.p, 256, C.byref(size)) == uc.UC_E_ARG
        finally:
            tiny.free()
        # just too large: refused when the first buffer that does not fit is carved
        too_big = uc.synth_kmers(14, 4_000_000, 31)
        rc, _, _ = noalloc(c, hip, too_big, 31, 11)
        assert rc == uc.UC_E_LIMIT
        assert b"scratch" in L.uc_last_error(c._ctx)
        # and the same call shape works again right after, on a job that fits
        rc, size, got = noalloc(c, hip, kmers, 31, 11)
        assert rc == uc.UC_OK and got == want
