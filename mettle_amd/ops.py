"""Operator calls on device tensors.

Thin wrappers that pass torch CUDA tensor addresses to libmettle_hip.so.
torch only provides device memory here; every computation runs in the HIP
library.  Each call synchronises torch's stream first (inputs made by torch
are then complete) and, for synchronous helpers, the library stream after.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import abi

_NP_TO_TORCH = {np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
                np.dtype(np.float32): torch.float32, np.dtype(np.float64): torch.float64,
                np.dtype(np.uint8): torch.uint8, np.dtype(np.uint32): torch.int32}


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the mettle HIP backend needs an MI355X")
    return torch.device("cuda:0")


def to_device(a: np.ndarray) -> torch.Tensor:
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.from_numpy(np.ascontiguousarray(a)).to(device())


def to_host(t: torch.Tensor, dtype=None) -> np.ndarray:
    a = t.detach().cpu().numpy()
    return a.view(dtype) if dtype is not None else a


def ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def sync() -> None:
    torch.cuda.synchronize()
    abi.check(abi.lib().mtl_stream_sync(None))


def _count(cnt: torch.Tensor, cap: int) -> int:
    """Device row count of a selection; a count above the capacity is a
    processing error, never a result."""
    n = int(cnt.item()) & 0xFFFFFFFF
    if n > cap:
        raise abi.MettleError(abi.MTL_ERR_PROCESSING, f"selection count {n} exceeds {cap}")
    return n


def _preds(preds):
    arr = (abi.Pred * max(1, len(preds)))()
    for i, (col, t, op, val) in enumerate(preds):
        arr[i] = abi.make_pred(ptr(col), t, op, val)
    return arr


def filter(preds, n_rows: int, in_sel: torch.Tensor | None = None,
           in_count: torch.Tensor | None = None):
    """preds: list of (device column, type, op, literal).  Returns (sel, count)."""
    torch.cuda.synchronize()
    cap = in_sel.numel() if in_sel is not None else n_rows
    out = torch.empty(max(cap, 4), dtype=torch.int32, device=device())
    cnt = torch.zeros(1, dtype=torch.int32, device=device())
    abi.check(abi.lib().mtl_filter(_preds(preds), len(preds), n_rows, ptr(in_sel),
                                   ptr(in_count), cap, ptr(out), ptr(cnt), None))
    sync()
    n = _count(cnt, cap)
    return out[:n], n


def filter_or(preds, or_next: int, n_rows: int):
    """mtl_filter_or: predicates p and p+1 OR-ed where bit p of or_next is
    set, the runs AND-ed.  Returns (sel, count)."""
    torch.cuda.synchronize()
    out = torch.empty(max(n_rows, 4), dtype=torch.int32, device=device())
    cnt = torch.zeros(1, dtype=torch.int32, device=device())
    abi.check(abi.lib().mtl_filter_or(_preds(preds), len(preds), or_next, n_rows, ptr(out),
                                      ptr(cnt), None))
    sync()
    n = _count(cnt, n_rows)
    return out[:n], n


def filter_project(preds, n_rows: int, col: torch.Tensor, t: int, op: int, value,
                   in_sel: torch.Tensor | None = None, in_count: torch.Tensor | None = None):
    """Fused filter + projection of the passing rows.

    Returns (sel, projected values, overflow flag); the values are
    col[sel[i]] op value, computed as the selection is written."""
    torch.cuda.synchronize()
    cap = in_sel.numel() if in_sel is not None else n_rows
    out = torch.empty(max(cap, 4), dtype=torch.int32, device=device())
    vals = torch.empty(max(cap, 4), dtype=col.dtype, device=device())
    cnt = torch.zeros(1, dtype=torch.int32, device=device())
    flag = torch.zeros(1, dtype=torch.int32, device=device())
    abi.check(abi.lib().mtl_filter_project(_preds(preds), len(preds), n_rows, ptr(in_sel),
                                           ptr(in_count), cap, ptr(col), t, op,
                                           abi.scalar_bits(value, t), ptr(out), ptr(cnt),
                                           ptr(vals), ptr(flag), None))
    sync()
    n = _count(cnt, cap)
    return out[:n], vals[:n], bool(flag.item())


def compare(preds, n_rows: int) -> torch.Tensor:
    torch.cuda.synchronize()
    out = torch.empty(max(n_rows, 1), dtype=torch.uint8, device=device())
    abi.check(abi.lib().mtl_compare(_preds(preds), len(preds), n_rows, ptr(out), None))
    sync()
    return out[:n_rows]


def gather(col: torch.Tensor, sel: torch.Tensor) -> torch.Tensor:
    torch.cuda.synchronize()
    out = torch.empty(max(sel.numel(), 1), dtype=col.dtype, device=device())
    abi.check(abi.lib().mtl_gather(ptr(col), col.element_size(), ptr(sel), None, sel.numel(),
                                   ptr(out), None))
    sync()
    return out[:sel.numel()]


def project(col: torch.Tensor, t: int, sel: torch.Tensor | None, op: int, value):
    torch.cuda.synchronize()
    n = sel.numel() if sel is not None else col.numel()
    out = torch.empty(max(n, 1), dtype=col.dtype, device=device())
    flag = torch.zeros(1, dtype=torch.int32, device=device())
    abi.check(abi.lib().mtl_project(ptr(col), t, ptr(sel), None, n, op,
                                    abi.scalar_bits(value, t), ptr(out), ptr(flag), None))
    sync()
    return out[:n], bool(flag.item())


def join(strategy: int, bkeys: torch.Tensor, btype: int, pkeys: torch.Tensor, ptype: int,
         min_key: int = 0, slots: int = 0, bsel: torch.Tensor | None = None,
         psel: torch.Tensor | None = None):
    """Build on bkeys, probe with pkeys; returns (probe rows, build rows)
    (JOIN_RADIX: in partition order)."""
    torch.cuda.synchronize()
    L = abi.lib()
    bn = bsel.numel() if bsel is not None else bkeys.numel()
    pn = psel.numel() if psel is not None else pkeys.numel()
    nbytes = L.mtl_join_table_bytes(strategy, bn, slots)
    table = torch.empty(max(int(nbytes), 16), dtype=torch.uint8, device=device())
    dup = torch.zeros(1, dtype=torch.int32, device=device())
    abi.check(L.mtl_join_build(strategy, ptr(bkeys), btype, ptr(bsel), None, bn, min_key,
                               slots, ptr(table), table.numel(), ptr(dup), None))
    sync()
    if int(dup.item()) & 2:
        raise abi.Unsupported(abi.MTL_UNSUPPORTED, "duplicate build key")
    if int(dup.item()) & 16:
        raise abi.Unsupported(abi.MTL_UNSUPPORTED, "radix partition over 3072 rows")
    if int(dup.item()) & 8:
        raise abi.MettleError(abi.MTL_ERR_INVALID_ARGUMENT, "build key outside the dense span")
    out_p = torch.empty(max(pn, 4), dtype=torch.int32, device=device())
    out_b = torch.empty(max(pn, 4), dtype=torch.int32, device=device())
    cnt = torch.zeros(1, dtype=torch.int32, device=device())
    probe_slots = slots if strategy == abi.JOIN_DENSE else bn
    abi.check(L.mtl_join_probe(strategy, ptr(table), table.numel(), min_key, probe_slots,
                               ptr(pkeys), ptype, ptr(psel), None, pn, ptr(out_p), ptr(out_b),
                               ptr(cnt), None))
    sync()
    n = int(cnt.item())
    return out_p[:n], out_b[:n]


class Groups:
    """Device-resident group table (mtl_groups handle)."""

    def __init__(self, handle, n_keys, aggs):
        self.h = handle
        self.n_keys = n_keys
        self.aggs = aggs  # list of (func, type)

    def release(self):
        if getattr(self, "h", None):
            abi.lib().mtl_groups_release(self.h)
            self.h = None

    def __del__(self):
        # at interpreter exit the module globals may already be gone
        if abi is not None and getattr(abi, "lib", None) is not None:
            self.release()

    def count(self) -> int:
        n = C.c_uint64()
        abi.check(abi.lib().mtl_groups_count(self.h, C.byref(n)))
        return n.value

    def _outs(self, cap):
        outs = [np.empty(max(cap, 1), dtype=abi.agg_result_dtype(f, t)) for f, t in self.aggs]
        ptrs = (C.c_void_p * max(1, len(outs)))(*[o.ctypes.data for o in outs])
        return outs, ptrs

    def fetch(self):
        cap = self.count()
        keys = np.empty((max(cap, 1), self.n_keys), dtype=np.int64)
        outs, ptrs = self._outs(cap)
        n = C.c_uint64()
        abi.check(abi.lib().mtl_groups_fetch(self.h, keys.ctypes.data, ptrs, cap, C.byref(n)))
        return keys[:n.value], [o[:n.value] for o in outs]

    def topk(self, agg: int, k: int, desc: bool = True):
        keys = np.empty((k, self.n_keys), dtype=np.int64)
        outs, ptrs = self._outs(k)
        n = C.c_uint32()
        abi.check(abi.lib().mtl_groups_topk(self.h, agg, 1 if desc else 0, k, keys.ctypes.data,
                                            ptrs, C.byref(n)))
        return keys[:n.value], [o[:n.value] for o in outs]


def group_agg(keys, aggs, n_rows: int, sel: torch.Tensor | None = None,
              sel_count: torch.Tensor | None = None) -> Groups:
    """keys: list of (device column, type, min, max); aggs: list of
    (func, device column | None, type[, DECIMAL scale])."""
    torch.cuda.synchronize()
    d = abi.make_group_desc(n_rows, [(ptr(c), t, mn, mx) for c, t, mn, mx in keys],
                            [(a[0], ptr(a[1])) + tuple(a[2:]) for a in aggs], sel=ptr(sel),
                            sel_count=ptr(sel_count),
                            sel_capacity=0 if sel is None else sel.numel())
    h = C.c_void_p()
    abi.check(abi.lib().mtl_group_agg(C.byref(d), C.byref(h), None))
    return Groups(h, len(keys), [(a[0], a[2]) for a in aggs])


def star_topk(desc: abi.StarDesc, stream=None):
    """Run the star pipeline synchronously; returns a list of StarRow."""
    torch.cuda.synchronize()
    out = (abi.StarRow * max(1, desc.k))()
    n = C.c_uint32(0)
    abi.check(abi.lib().mtl_star_topk(C.byref(desc), out, C.byref(n), stream))
    return [out[i] for i in range(n.value)]


# ---------------------------------------------------------------------------
# Serving loop
# ---------------------------------------------------------------------------
def scan_query(preds, n_rows: int, proj=None) -> abi.Query:
    """preds: list of (device column, type, op, literal); proj: optional
    (device column, type, op, scalar)."""
    q = abi.Query()
    q.kind = abi.Q_SCAN
    q.n_rows = n_rows
    q.n_preds = len(preds)
    for i, (c, t, op, lit) in enumerate(preds):
        q.preds[i] = abi.make_pred(ptr(c), t, op, lit)
    if proj is not None:
        c, t, op, sc = proj
        q.proj_col = C.c_void_p(ptr(c))
        q.proj_type = t
        q.proj_op = op
        q.proj_scalar = abi.scalar_bits(sc, t)
    return q


def star_query(desc: abi.StarDesc):
    """Returns (query, output rows buffer); keep the buffer alive until the
    query completes."""
    q = abi.Query()
    q.kind = abi.Q_STAR
    q.star = desc
    out = (abi.StarRow * max(1, desc.k))()
    q.star_out = C.cast(out, C.POINTER(abi.StarRow))
    return q, out


def group_query(keys, aggs, n_rows: int, k: int = 0, agg: int = 0, desc: bool = True,
                capacity: int = 0, sel: torch.Tensor | None = None):
    """A grouped-aggregate query (keys, aggs as for group_agg): the top k
    groups by aggregate `agg`, or every group (at most `capacity`) when k is
    0.  Returns (query, (keys array, [per-aggregate arrays])); keep both
    alive until the query completes."""
    q = abi.Query()
    q.kind = abi.Q_GROUP
    q.group = abi.make_group_desc(n_rows, [(ptr(c), t, mn, mx) for c, t, mn, mx in keys],
                                  [(a[0], ptr(a[1])) + tuple(a[2:]) for a in aggs],
                                  sel=ptr(sel), sel_capacity=0 if sel is None else sel.numel())
    cap = k if k else capacity
    q.group_k = k
    q.group_agg = agg
    q.group_desc = 1 if desc else 0
    q.group_capacity = cap
    kout = np.zeros((max(cap, 1), len(keys)), dtype=np.int64)
    outs = [np.zeros(max(cap, 1), dtype=abi.agg_result_dtype(a[0], a[2])) for a in aggs]
    q.group_keys = kout.ctypes.data
    for i, o in enumerate(outs):
        q.group_out[i] = o.ctypes.data
    return q, (kout, outs)


class Server:
    """mtl_server: n_slots worker streams; each worker runs the queries that
    have arrived in batches of up to max_batch (shared scans), and
    n_slots = max_batch = 1 is serial dispatch."""

    def __init__(self, n_slots: int, budget: int = 0, max_batch: int = 1):
        self.h = C.c_void_p()
        abi.check(abi.lib().mtl_server_create(n_slots, max_batch, budget, C.byref(self.h)))

    def submit(self, q: abi.Query, not_before_ns: int = 0) -> int:
        t = C.c_uint64()
        abi.check(abi.lib().mtl_server_submit(self.h, C.byref(q), not_before_ns, C.byref(t)))
        return t.value

    def submit_sql(self, sql: str, not_before_ns: int = 0) -> int:
        t = C.c_uint64()
        abi.check(abi.lib().mtl_server_submit_sql(self.h, sql.encode(), not_before_ns,
                                                  C.byref(t)))
        return t.value

    def wait(self, ticket: int) -> abi.QueryResult:
        """The result record; raises MettleError if the query failed."""
        r = abi.QueryResult()
        st = abi.lib().mtl_server_wait(self.h, ticket, C.byref(r))
        msg = r.error.decode(errors="replace")
        if r.status == abi.MTL_UNSUPPORTED:
            raise abi.Unsupported(r.status, msg)
        if r.status != abi.MTL_OK:
            raise abi.MettleError(r.status, msg)
        abi.check(st)
        return r

    def poll(self, ticket: int):
        """The result record if the query has completed (the ticket is then
        released), else None."""
        r = abi.QueryResult()
        done = C.c_int32()
        st = abi.lib().mtl_server_poll(self.h, ticket, C.byref(r), C.byref(done))
        if not done.value:
            abi.check(st)
            return None
        if r.status != abi.MTL_OK:
            raise abi.MettleError(r.status, r.error.decode(errors="replace"))
        abi.check(st)
        return r

    def wait_sql(self, ticket: int):
        """A SQL query's rows, as mettle_amd.sql.execute returns them."""
        from mettle_amd import sql as S
        r = self.wait(ticket)
        return S.take_result(r.result)

    def close(self):
        if self.h:
            abi.check(abi.lib().mtl_server_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def sort(column: torch.Tensor, type_code: int, n: int, sel: torch.Tensor | None = None,
         count: torch.Tensor | None = None, desc: bool = False):
    """ORDER BY column over the selection (or rows 0..n-1); returns (device
    row ids, count)."""
    torch.cuda.synchronize()
    cap = n if sel is None else sel.numel()
    out = torch.empty(max(cap, 1), dtype=torch.int32, device="cuda")
    got = C.c_uint32()
    abi.check(abi.lib().mtl_sort(ptr(column), type_code, ptr(sel), ptr(count), cap,
                                 1 if desc else 0, ptr(out), C.byref(got), None))
    torch.cuda.synchronize()
    return out[:got.value], got.value


# ---------------------------------------------------------------------------
# Tuned launch parameters
# ---------------------------------------------------------------------------
TUNE_SMALL_SCAN, TUNE_JOIN, TUNE_GROUP, TUNE_ALL = 1, 2, 4, 7


def tune_describe() -> dict:
    """The tuned parameter table as {name: value}."""
    buf = C.create_string_buffer(512)
    abi.check(abi.lib().mtl_tune_describe(buf, len(buf)))
    return {k: int(v) for k, v in (kv.split("=") for kv in buf.value.decode().split())}


def tune_set(name: str, value: int) -> None:
    abi.check(abi.lib().mtl_tune_set(name.encode(), value))


def autotune(groups: int = TUNE_ALL) -> str:
    """Run the autotuner on the device; returns its report."""
    buf = C.create_string_buffer(16384)
    abi.check(abi.lib().mtl_autotune(groups, buf, len(buf)))
    return buf.value.decode()
