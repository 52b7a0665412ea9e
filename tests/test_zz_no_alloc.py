"""mtl_filter_noalloc (the filter whose scratch is a static device array)
against mtl_filter through the C ABI: two calls in a row, the second with
a smaller input, dense and through an input selection, and a request over
its row limit refused before any GPU work."""
import ctypes as C

import numpy as np
import pytest

from mettle_amd import abi

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _run(L, fn, preds, n, in_sel, cap, stream):
    out = torch.empty(max(cap, 4), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    isel = None if in_sel is None else in_sel.data_ptr()
    torch.cuda.synchronize()
    abi.check(fn(preds, len(preds), n, isel, None, cap if in_sel is not None else 0,
                 out.data_ptr(), cnt.data_ptr(), stream))
    abi.check(L.mtl_stream_sync(stream))
    k = int(cnt.item())
    return out[:k].cpu().numpy().view(np.uint32)


def test_noalloc_filter_matches_filter():
    L = abi.lib()
    s = C.c_void_p()
    abi.check(L.mtl_stream_create(C.byref(s)))
    try:
        rng = np.random.default_rng(70)
        for n, frac in ((3_000_000, 0.3), (70_001, 0.6)):  # second call smaller
            a = rng.integers(-10**6, 10**6, n).astype(np.int32)
            b = rng.standard_normal(n).astype(np.float32)
            da = torch.from_numpy(a).cuda()
            db = torch.from_numpy(b).cuda()
            lit = int(-10**6 + frac * 2 * 10**6)
            preds = (abi.Pred * 2)(abi.make_pred(da.data_ptr(), abi.INTEGER, abi.CMP_LT, lit),
                                   abi.make_pred(db.data_ptr(), abi.FLOAT, abi.CMP_GE, -0.5))
            ref = _run(L, L.mtl_filter, preds, n, None, n, s)
            got = _run(L, L.mtl_filter_noalloc, preds, n, None, n, s)
            np.testing.assert_array_equal(got, ref)
            exp = np.flatnonzero((a < lit) & (b >= np.float32(-0.5))).astype(np.uint32)
            np.testing.assert_array_equal(got, exp)
            # through an input selection (every third row)
            sel = torch.from_numpy(np.arange(0, n, 3, dtype=np.int32)).cuda()
            ref = _run(L, L.mtl_filter, preds, n, sel, sel.numel(), s)
            got = _run(L, L.mtl_filter_noalloc, preds, n, sel, sel.numel(), s)
            np.testing.assert_array_equal(got, ref)
        # over the limit: refused on the host
        preds1 = (abi.Pred * 1)(abi.make_pred(da.data_ptr(), abi.INTEGER, abi.CMP_LT, 0))
        st = L.mtl_filter_noalloc(preds1, 1, 100_000_001, None, None, 0, da.data_ptr(),
                                  da.data_ptr(), s)
        assert st == abi.MTL_ERR_INVALID_ARGUMENT
        assert "at most 100000000 rows" in abi.last_error()
    finally:
        L.mtl_stream_destroy(s)
