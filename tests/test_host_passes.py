"""Host helpers of the setup path (no GPU): slb_par_for splits a pass into
contiguous ranges in thread order that cover [0, n) exactly once, and
slb_big_calloc returns zero-filled, 2 MiB-aligned node arrays (small ones
from calloc) that slb_big_free releases."""
import ctypes as C
import threading

import numpy as np
import pytest

RANGE_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_long, C.c_long, C.c_int)


@pytest.fixture(scope="module")
def lib():
    import sunlightlb_amd as s
    L = s.lib()
    L.slb_par_for.restype = C.c_int
    L.slb_par_for.argtypes = [C.c_long, RANGE_FN, C.c_void_p]
    L.slb_host_threads.restype = C.c_int
    L.slb_big_calloc.restype = C.c_void_p
    L.slb_big_calloc.argtypes = [C.c_long, C.c_size_t]
    L.slb_big_free.restype = None
    L.slb_big_free.argtypes = [C.c_void_p]
    return L


@pytest.mark.parametrize("n", [0, 1, 1000, 1 << 18, (1 << 18) + 7, 3_000_001])
def test_par_for_ranges_cover_once_in_order(lib, n):
    seen = []
    lock = threading.Lock()

    def fn(ctx, lo, hi, tid):
        with lock:
            seen.append((tid, lo, hi))

    cb = RANGE_FN(fn)
    nr = lib.slb_par_for(n, cb, None)
    seen.sort()
    assert [t for t, _, _ in seen] == list(range(nr))
    assert 1 <= nr <= max(1, lib.slb_host_threads())
    if n < (1 << 18):
        assert nr == 1
    # contiguous, in thread order, covering [0, n)
    edge = 0
    for _, lo, hi in seen:
        assert lo == edge and hi >= lo
        edge = hi
    assert edge == n


@pytest.mark.parametrize("count, size", [(1, 8), (1000, 4), (1 << 20, 8), (3 << 20, 4)])
def test_big_calloc_zero_filled_and_aligned(lib, count, size):
    p = lib.slb_big_calloc(count, size)
    assert p
    a = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(count * size,))
    assert not a.any()
    if count * size >= 2 << 20:
        assert p % (2 << 20) == 0
    a[:] = 0xAB                     # writable over its whole length
    lib.slb_big_free(p)
    lib.slb_big_free(None)          # a no-op, as free(NULL)
