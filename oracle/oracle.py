"""Python access to the CPU oracle (oracle/mettle_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the source of expected results.  Nothing under
mettle_amd/ imports this module.

All pointers handed to the oracle are host (numpy) pointers.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from mettle_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libmettle_oracle.so")
_lib = None


def build() -> None:
    subprocess.run(["make", "-B", "-C", _HERE], check=True, stdout=subprocess.DEVNULL)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, u64, i32, i64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_int64
        L.orc_pred_row.restype = C.c_int
        L.orc_pred_row.argtypes = [C.POINTER(abi.Pred), u64]
        L.orc_filter.restype = u64
        L.orc_filter.argtypes = [C.POINTER(abi.Pred), u32, u64, vp, u64, vp]
        L.orc_project.restype = i32
        L.orc_project.argtypes = [vp, i32, vp, u64, i32, i64, vp, C.POINTER(u32)]
        L.orc_join.restype = i64
        L.orc_join.argtypes = [vp, i32, vp, u64, vp, i32, vp, u64, vp, vp]
        L.orc_star_topk.restype = i32
        L.orc_star_topk.argtypes = [C.POINTER(abi.StarDesc), C.POINTER(abi.StarRow),
                                    C.POINTER(u32)]
        L.orc_date_days.restype = i32
        L.orc_date_days.argtypes = [i32, i32, i32]
        L.orc_group_agg.restype = i64
        L.orc_group_agg.argtypes = [C.POINTER(abi.GroupDesc), vp, C.POINTER(vp), u64]
        _lib = L
    return _lib


def _ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


def preds_array(preds):
    arr = (abi.Pred * max(1, len(preds)))()
    for i, (col, t, op, val) in enumerate(preds):
        arr[i] = abi.make_pred(_ptr(col), t, op, val)
    return arr


def filter(preds, n_rows: int, in_sel: np.ndarray | None = None) -> np.ndarray:
    """preds: list of (np column, type, op, literal). Returns uint32 row ids."""
    n = len(in_sel) if in_sel is not None else n_rows
    out = np.empty(max(n, 1), dtype=np.uint32)
    arr = preds_array(preds)
    cnt = lib().orc_filter(arr, len(preds), n_rows, _ptr(in_sel), n, _ptr(out))
    return out[:cnt].copy()


def project(col: np.ndarray, t: int, sel: np.ndarray | None, op: int, value):
    n = len(sel) if sel is not None else len(col)
    out = np.empty(max(n, 1), dtype=col.dtype)
    ovf = C.c_uint32(0)
    lib().orc_project(_ptr(col), t, _ptr(sel), n, op, abi.scalar_bits(value, t), _ptr(out),
                      C.byref(ovf))
    return out[:n].copy(), bool(ovf.value)


def join(bkeys, btype, pkeys, ptype, bsel=None, psel=None):
    bn = len(bsel) if bsel is not None else len(bkeys)
    pn = len(psel) if psel is not None else len(pkeys)
    op = np.empty(max(pn, 1), dtype=np.uint32)
    ob = np.empty(max(pn, 1), dtype=np.uint32)
    n = lib().orc_join(_ptr(bkeys), btype, _ptr(bsel), bn, _ptr(pkeys), ptype, _ptr(psel), pn,
                       _ptr(op), _ptr(ob))
    if n == -1:
        raise abi.Unsupported(abi.MTL_UNSUPPORTED, "duplicate build key")
    if n < 0:
        raise ValueError("join span too large for the oracle")
    return op[:n].copy(), ob[:n].copy()


def star_topk(desc: abi.StarDesc):
    """desc must hold host pointers. Returns (status, list of StarRow)."""
    out = (abi.StarRow * max(1, desc.k))()
    n = C.c_uint32(0)
    st = lib().orc_star_topk(C.byref(desc), out, C.byref(n))
    return st, [out[i] for i in range(n.value)]


def group_agg(keys, aggs, sel=None):
    """keys: list of (np column, type); aggs: list of (func, np column | None,
    type[, DECIMAL scale]).  Returns (keys int64 [n, n_keys], [per-agg numpy
    arrays]) in ascending key order."""
    n_rows = len(keys[0][0])
    kk = [(k.ctypes.data, t, int(k.min()) if len(k) else 0, int(k.max()) if len(k) else 0)
          for k, t in keys]
    aa = [(a[0], None if a[1] is None else a[1].ctypes.data) + tuple(a[2:]) for a in aggs]
    d = abi.make_group_desc(n_rows, kk, aa, sel=None if sel is None else sel.ctypes.data,
                            sel_capacity=0 if sel is None else len(sel))
    cap = max(1, len(sel) if sel is not None else n_rows)
    kout = np.empty((cap, len(keys)), dtype=np.int64)
    outs = [np.empty(cap, dtype=abi.agg_result_dtype(a[0], a[2])) for a in aggs]
    ptrs = (C.c_void_p * max(1, len(outs)))(*[o.ctypes.data for o in outs])
    n = lib().orc_group_agg(C.byref(d), kout.ctypes.data, ptrs, cap)
    return kout[:n].copy(), [o[:n].copy() for o in outs]


def date_days(y: int, m: int, d: int) -> int:
    return lib().orc_date_days(y, m, d)
