"""Refcounted handles of the boundary (GpuBuffer / GpuVector,
gpu_buffer.rs:12-52) and the serving loop's non-blocking poll: memory is
charged to the processing budget and returned at the last release; a
selection vector's count is read with mtl_vec_count; mtl_filter_vec equals
the oracle's filter, also over an input vector; mtl_server_poll returns the
same record mtl_server_wait would."""
import ctypes as C
import time

import numpy as np
import pytest

from mettle_amd import abi
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture
def lib():
    L = abi.lib()
    # the budget is counted from what earlier work in this process still
    # holds (stream arenas are kept for reuse), plus 4 GiB for these tests
    held = C.c_uint64()
    abi.check(L.mtl_processing_in_use(C.byref(held)))
    abi.check(L.mtl_init(1 << 30, held.value + (4 << 30), 0, None, 0))
    yield L
    L.mtl_cache_reset()
    L.mtl_shutdown()


def _in_use(L):
    b = C.c_uint64()
    abi.check(L.mtl_processing_in_use(C.byref(b)))
    return b.value


def test_buffer_refcount_and_budget(lib):
    L = lib
    base = _in_use(L)
    b = C.c_void_p()
    abi.check(L.mtl_buf_alloc(1 << 20, C.byref(b)))
    assert L.mtl_buf_bytes(b) == 1 << 20 and _in_use(L) >= base + (1 << 20)
    data = np.arange(1 << 18, dtype=np.uint32)
    abi.check(L.mtl_memcpy_h2d(L.mtl_buf_device(b), data.ctypes.data, data.nbytes, None))
    abi.check(L.mtl_buf_retain(b))
    abi.check(L.mtl_buf_release(b))  # one reference left: still readable
    back = np.empty_like(data)
    abi.check(L.mtl_memcpy_d2h(back.ctypes.data, L.mtl_buf_device(b), data.nbytes, None))
    np.testing.assert_array_equal(back, data)
    abi.check(L.mtl_buf_release(b))
    assert _in_use(L) == base
    with pytest.raises(abi.MettleError):
        abi.check(L.mtl_buf_alloc(16 << 30, C.byref(b)))  # beyond the 4 GiB budget


@pytest.mark.parametrize("n", [0, 5000, 1_000_003])
def test_filter_vec_matches_the_oracle(lib, n):
    L = lib
    rng = np.random.default_rng(n + 1)
    a = rng.integers(-1000, 1000, max(n, 1)).astype(np.int32)[:n]
    b = rng.integers(0, 10, max(n, 1)).astype(np.int64)[:n]
    da = torch.from_numpy(a if n else np.zeros(1, np.int32)).cuda()
    db = torch.from_numpy(b if n else np.zeros(1, np.int64)).cuda()
    v1, v2 = C.c_void_p(), C.c_void_p()
    abi.check(L.mtl_vec_alloc(n, C.byref(v1)))
    abi.check(L.mtl_vec_alloc(n, C.byref(v2)))
    try:
        p1 = (abi.Pred * 1)(abi.make_pred(da.data_ptr(), abi.INTEGER, abi.CMP_LT, 100))
        abi.check(L.mtl_filter_vec(p1, 1, n, None, v1, None))
        cnt = C.c_uint64()
        abi.check(L.mtl_vec_count(v1, None, C.byref(cnt)))
        ref1 = orc.filter([(a, abi.INTEGER, abi.CMP_LT, 100)], n)
        assert cnt.value == len(ref1) and L.mtl_vec_capacity(v1) == n
        got = np.empty(max(cnt.value, 1), np.uint32)
        if cnt.value:
            abi.check(L.mtl_memcpy_d2h(got.ctypes.data, L.mtl_vec_rows(v1), cnt.value * 4, None))
        np.testing.assert_array_equal(got[:cnt.value], ref1)
        # refine through the first vector
        p2 = (abi.Pred * 1)(abi.make_pred(db.data_ptr(), abi.BIGINT, abi.CMP_EQ, 3))
        abi.check(L.mtl_filter_vec(p2, 1, n, v1, v2, None))
        abi.check(L.mtl_vec_count(v2, None, C.byref(cnt)))
        ref2 = orc.filter([(b, abi.BIGINT, abi.CMP_EQ, 3)], n, in_sel=ref1)
        got = np.empty(max(cnt.value, 1), np.uint32)
        if cnt.value:
            abi.check(L.mtl_memcpy_d2h(got.ctypes.data, L.mtl_vec_rows(v2), cnt.value * 4, None))
        np.testing.assert_array_equal(got[:cnt.value], ref2)
        # an output vector smaller than the input is refused
        small = C.c_void_p()
        abi.check(L.mtl_vec_alloc(max(n - 1, 0), C.byref(small)))
        if n:
            with pytest.raises(abi.MettleError):
                abi.check(L.mtl_filter_vec(p1, 1, n, None, small, None))
        abi.check(L.mtl_vec_release(small))
    finally:
        L.mtl_vec_release(v1)
        L.mtl_vec_release(v2)


def test_server_poll(lib):
    from mettle_amd import ops
    rng = np.random.default_rng(3)
    n = 2_000_000
    v = rng.integers(0, 1 << 20, n).astype(np.int32)
    dv = ops.to_device(v)
    q = ops.scan_query([(dv, abi.INTEGER, abi.CMP_GT, 1 << 19)], n,
                       (dv, abi.INTEGER, abi.PROJ_MUL, 3))
    with ops.Server(1, max_batch=4) as srv:
        t_wait = srv.submit(q)
        r_wait = srv.wait(t_wait)
        t = srv.submit(q)
        r = None
        deadline = time.time() + 30
        while r is None and time.time() < deadline:
            r = srv.poll(t)
            if r is None:
                time.sleep(0.0005)
        assert r is not None
        assert (r.count, r.fingerprint) == (r_wait.count, r_wait.fingerprint)
        assert r.count == int((v > (1 << 19)).sum())
        with pytest.raises(abi.MettleError):  # the ticket was released
            srv.poll(t)
