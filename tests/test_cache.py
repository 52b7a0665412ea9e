"""HBM column cache: uploads through the pinned chunks are exact, repeated
puts hit, versions re-upload, the LRU evicts unreferenced columns within
the cache budget, and referenced columns are never evicted."""
import ctypes as C

import numpy as np
import pytest

from mettle_amd import abi

pytestmark = pytest.mark.gpu

MB = 1 << 20


@pytest.fixture
def lib():
    L = abi.lib()
    yield L
    L.mtl_cache_reset()
    L.mtl_shutdown()


def _init(L, cache_bytes, processing_bytes=256 * MB):
    abi.check(L.mtl_init(cache_bytes, processing_bytes, 0, None, 0))


def _put(L, t, c, arr, version=0):
    p = C.c_void_p()
    abi.check(L.mtl_cache_put(t, c, arr.ctypes.data, arr.nbytes, version, C.byref(p)))
    return p.value


def _read(L, dev, nbytes, dtype):
    out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
    abi.check(L.mtl_memcpy_d2h(out.ctypes.data, dev, nbytes, None))
    return out


def _stats(L):
    v = [C.c_uint64() for _ in range(6)]
    abi.check(L.mtl_cache_stats(*[C.byref(x) for x in v]))
    return dict(zip(["entries", "bytes", "hits", "misses", "evictions", "uploaded"],
                    [x.value for x in v]))


def test_put_roundtrip_hit_and_version(lib):
    _init(lib, 512 * MB)
    rng = np.random.default_rng(1)
    a = rng.integers(-2**31, 2**31, 25_000_001, dtype=np.int64).astype(np.int32)  # 4 chunks
    p = _put(lib, 7, 0, a)
    np.testing.assert_array_equal(_read(lib, p, a.nbytes, np.int32), a)
    assert _put(lib, 7, 0, a) == p  # hit: same device copy
    s = _stats(lib)
    assert (s["entries"], s["hits"], s["misses"], s["uploaded"]) == (1, 1, 1, a.nbytes)
    abi.check(lib.mtl_cache_release(7, 0))
    abi.check(lib.mtl_cache_release(7, 0))
    b = a[::-1].copy()
    p2 = _put(lib, 7, 0, b, version=1)  # new version: uploaded again
    np.testing.assert_array_equal(_read(lib, p2, b.nbytes, np.int32), b)
    abi.check(lib.mtl_cache_release(7, 0))


def test_lru_eviction_and_references(lib):
    _init(lib, 64 * MB)
    cols = [np.full(6 * MB, i, dtype=np.int32) for i in range(4)]  # 24 MiB each
    _put(lib, 1, 0, cols[0])
    _put(lib, 1, 1, cols[1])
    abi.check(lib.mtl_cache_release(1, 0))
    abi.check(lib.mtl_cache_release(1, 1))
    got = C.c_void_p()
    abi.check(lib.mtl_cache_get(1, 0, C.byref(got), None))  # column 0 becomes most recent
    abi.check(lib.mtl_cache_release(1, 0))
    p2 = _put(lib, 1, 2, cols[2])  # evicts column 1, the least recently used
    assert _stats(lib)["evictions"] == 1
    with pytest.raises(abi.MettleError):
        abi.check(lib.mtl_cache_get(1, 1, C.byref(got), None))
    np.testing.assert_array_equal(_read(lib, p2, cols[2].nbytes, np.int32), cols[2])
    # columns 0 (free) and 2 (referenced): a fourth column evicts 0, not 2
    _put(lib, 1, 3, cols[3])
    assert _stats(lib)["evictions"] == 2
    # both remaining columns are referenced: no room for another
    with pytest.raises(abi.MettleError) as ei:
        _put(lib, 2, 0, cols[0])
    assert ei.value.code == abi.MTL_ERR_BUDGET
    with pytest.raises(abi.MettleError):
        abi.check(lib.mtl_cache_drop(1))  # referenced
    abi.check(lib.mtl_cache_release(1, 2))
    abi.check(lib.mtl_cache_release(1, 3))
    abi.check(lib.mtl_cache_drop(1))
    assert _stats(lib)["entries"] == 0


def test_put_requires_init(lib):
    lib.mtl_shutdown()
    a = np.zeros(16, dtype=np.int32)
    p = C.c_void_p()
    assert lib.mtl_cache_put(1, 0, a.ctypes.data, a.nbytes, 0, C.byref(p)) == \
        abi.MTL_ERR_NOT_INITIALIZED
