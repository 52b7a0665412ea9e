"""ctypes view of the C ABI declared in include/mettle_hip.h.

This module only loads ``libmettle_hip.so`` (built in-tree by
``__graft_entry__.build()``) and declares its structs and prototypes.  There
is no fallback: if the library is missing, :func:`lib` raises.
"""
from __future__ import annotations

import ctypes as C
import os
import struct

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmettle_hip.so")

# status codes (MettleError, src/error.rs:5-11, + fallback/overflow)
MTL_OK = 0
MTL_ERR_INVALID_BUFFER_CONFIG = 1
MTL_ERR_DEVICE_UNAVAILABLE = 2
MTL_ERR_NOT_INITIALIZED = 3
MTL_ERR_PROCESSING = 4
MTL_UNSUPPORTED = 5
MTL_ERR_INVALID_ARGUMENT = 6
MTL_ERR_OVERFLOW = 7
MTL_ERR_BUDGET = 8

# logical types
INTEGER, BIGINT, FLOAT, DOUBLE, DATE, TIMESTAMP, DECIMAL, DICT8, DICT32, HUGEINT, VARCHAR = range(1, 12)
TYPE_WIDTH = {INTEGER: 4, BIGINT: 8, FLOAT: 4, DOUBLE: 8, DATE: 4, TIMESTAMP: 8,
              DECIMAL: 8, DICT8: 1, DICT32: 4, HUGEINT: 16}
TYPE_NAME = {INTEGER: "INTEGER", BIGINT: "BIGINT", FLOAT: "FLOAT", DOUBLE: "DOUBLE",
             DATE: "DATE", TIMESTAMP: "TIMESTAMP", DECIMAL: "DECIMAL", DICT8: "DICT8",
             DICT32: "DICT32", HUGEINT: "HUGEINT", VARCHAR: "VARCHAR"}

# compare ops (ScalarCompareOp codes, src/metal/shaders/mod.rs:27-36)
CMP_EQ, CMP_LT, CMP_LE, CMP_GT, CMP_GE = range(5)
CMP_SYMBOL = {"=": CMP_EQ, "<": CMP_LT, "<=": CMP_LE, ">": CMP_GT, ">=": CMP_GE}
# projection ops (ScalarProjectOp codes, src/metal/shaders/mod.rs:16-25)
PROJ_ADD, PROJ_MUL, PROJ_SUB, PROJ_SUBFROM = range(4)
# expression program ops
EXPR_COL, EXPR_CONST, EXPR_ADD, EXPR_SUB, EXPR_MUL = range(5)
EXPR_MAX_OPS = 16
EXPR_MAX_COLS = 4
MAX_PREDS = 8
STAR_MAX_TIES = 3
# join strategies
JOIN_DENSE, JOIN_HASH, JOIN_RADIX = 0, 1, 2


class Pred(C.Structure):
    _fields_ = [("column", C.c_void_p), ("type", C.c_int32), ("op", C.c_int32),
                ("scalar", C.c_int64)]


class Expr(C.Structure):
    _fields_ = [("n_ops", C.c_uint32), ("is_double", C.c_int32),
                ("op", C.c_int32 * EXPR_MAX_OPS), ("arg", C.c_int64 * EXPR_MAX_OPS),
                ("n_cols", C.c_uint32), ("col", C.c_void_p * EXPR_MAX_COLS),
                ("col_type", C.c_int32 * EXPR_MAX_COLS)]


class StarDesc(C.Structure):
    _fields_ = [
        ("d_rows", C.c_uint64), ("d_key", C.c_void_p), ("d_key_type", C.c_int32),
        ("d_key_min", C.c_int64), ("d_key_max", C.c_int64), ("d_n_preds", C.c_uint32),
        ("d_preds", Pred * MAX_PREDS),
        ("p_rows", C.c_uint64), ("p_key", C.c_void_p), ("p_key_type", C.c_int32),
        ("p_key_min", C.c_int64), ("p_key_max", C.c_int64), ("p_fk", C.c_void_p),
        ("p_fk_type", C.c_int32), ("p_n_preds", C.c_uint32), ("p_preds", Pred * MAX_PREDS),
        ("f_rows", C.c_uint64), ("f_fk", C.c_void_p), ("f_fk_type", C.c_int32),
        ("f_n_preds", C.c_uint32), ("f_preds", Pred * MAX_PREDS),
        ("agg", Expr),
        ("k", C.c_uint32), ("n_ties", C.c_uint32),
        ("tie_col", C.c_void_p * STAR_MAX_TIES), ("tie_type", C.c_int32 * STAR_MAX_TIES),
        ("tie_desc", C.c_int32 * STAR_MAX_TIES),
    ]


GROUP_MAX_KEYS = 4
GROUP_MAX_AGGS = 8
AGG_SUM, AGG_COUNT, AGG_MIN, AGG_MAX, AGG_AVG = range(5)


class AggSpec(C.Structure):
    _fields_ = [("func", C.c_int32), ("type", C.c_int32), ("values", C.c_void_p),
                ("scale", C.c_int32)]


class GroupDesc(C.Structure):
    _fields_ = [
        ("n_rows", C.c_uint64), ("sel", C.c_void_p), ("sel_count", C.c_void_p),
        ("sel_capacity", C.c_uint64), ("n_keys", C.c_uint32),
        ("key_col", C.c_void_p * GROUP_MAX_KEYS), ("key_type", C.c_int32 * GROUP_MAX_KEYS),
        ("key_min", C.c_int64 * GROUP_MAX_KEYS), ("key_max", C.c_int64 * GROUP_MAX_KEYS),
        ("n_aggs", C.c_uint32), ("aggs", AggSpec * GROUP_MAX_AGGS),
    ]


def agg_result_dtype(func: int, type_code: int):
    """numpy dtype of one aggregate's per-group output (see mtl_groups_fetch)."""
    import numpy as np
    is_f = type_code in (FLOAT, DOUBLE)
    if func == AGG_COUNT:
        return np.dtype(np.int64)
    if func == AGG_SUM:
        return np.dtype(np.float64) if is_f else np.dtype([("lower", "<u8"), ("upper", "<i8")])
    if func == AGG_AVG:
        return np.dtype(np.float64)
    return np.dtype(np.float64) if is_f else np.dtype(np.int64)


def make_group_desc(n_rows, keys, aggs, sel=None, sel_count=None, sel_capacity=0):
    """keys: list of (ptr, type, min, max); aggs: list of (func, ptr, type)
    or (func, ptr, type, scale) -- scale: DECIMAL digits after the point."""
    d = GroupDesc()
    d.n_rows = n_rows
    d.sel = sel
    d.sel_count = sel_count
    d.sel_capacity = sel_capacity
    d.n_keys = len(keys)
    for i, (p, t, mn, mx) in enumerate(keys):
        d.key_col[i] = p
        d.key_type[i] = t
        d.key_min[i] = int(mn)
        d.key_max[i] = int(mx)
    d.n_aggs = len(aggs)
    for i, a in enumerate(aggs):
        f, p, t = a[:3]
        d.aggs[i] = AggSpec(f, t, p, a[3] if len(a) > 3 else 0)
    return d


def hugeint_to_int(rec) -> int:
    return (int(rec["upper"]) << 64) | int(rec["lower"])


class StarRow(C.Structure):
    _fields_ = [("p_row", C.c_uint32), ("count", C.c_uint32), ("sum_i64", C.c_int64),
                ("sum_f64", C.c_double)]


Q_SCAN = 0
Q_STAR = 1
Q_GROUP = 2
Q_SQL = 3
HOOK_ADD, HOOK_MUL, HOOK_EQ, HOOK_LT, HOOK_LE, HOOK_GT, HOOK_GE = range(7)


class Query(C.Structure):
    _fields_ = [("kind", C.c_int32), ("n_rows", C.c_uint64), ("n_preds", C.c_uint32),
                ("preds", Pred * MAX_PREDS), ("proj_col", C.c_void_p), ("proj_type", C.c_int32),
                ("proj_op", C.c_int32), ("proj_scalar", C.c_int64), ("star", StarDesc),
                ("star_out", C.POINTER(StarRow)),
                ("group", GroupDesc), ("group_agg", C.c_uint32), ("group_desc", C.c_int32),
                ("group_k", C.c_uint32), ("group_capacity", C.c_uint64),
                ("group_keys", C.c_void_p), ("group_out", C.c_void_p * GROUP_MAX_AGGS),
                ("sql", C.c_char_p)]


class QueryResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("n_rows", C.c_uint32), ("count", C.c_uint64),
                ("fingerprint", C.c_uint64), ("t_submit_ns", C.c_int64),
                ("t_start_ns", C.c_int64), ("t_done_ns", C.c_int64),
                ("error", C.c_char * 160), ("groups", C.c_uint64), ("result", C.c_void_p),
                ("batch", C.c_uint32)]


def scalar_bits(value, type_code: int) -> int:
    """Literal -> the int64 bit pattern the ABI expects for that column type."""
    if type_code in (FLOAT, DOUBLE):
        return struct.unpack("<q", struct.pack("<d", float(value)))[0]
    return int(value)


def make_pred(column_ptr: int, type_code: int, op: int, value) -> Pred:
    return Pred(C.c_void_p(column_ptr), type_code, op, scalar_bits(value, type_code))


def make_expr(program, cols, is_double: bool = False) -> Expr:
    """program: list of (op, arg) in postfix; cols: list of (ptr, type)."""
    e = Expr()
    if len(program) > EXPR_MAX_OPS or len(cols) > EXPR_MAX_COLS:
        raise ValueError("expression too large")
    e.n_ops = len(program)
    e.is_double = 1 if is_double else 0
    for i, (op, arg) in enumerate(program):
        e.op[i] = op
        if op == EXPR_CONST and is_double:
            arg = scalar_bits(arg, DOUBLE)
        e.arg[i] = int(arg)
    e.n_cols = len(cols)
    for i, (ptr, t) in enumerate(cols):
        e.col[i] = C.c_void_p(ptr)
        e.col_type[i] = t
    return e


_lib = None


def lib() -> C.CDLL:
    """Load libmettle_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: run __graft_entry__.build() (make -C mettle_amd)")
    # torch bundles its own HIP runtime. When this library's runtime was loaded
    # first and torch's second, this library's hipGetDeviceCount reported no
    # device on the MI355X box; loading torch first gives the order in which
    # both runtimes see the GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    i32, u32, u64, i64, vp = C.c_int32, C.c_uint32, C.c_uint64, C.c_int64, C.c_void_p
    sig = {
        "mtl_parse_bytes": (i32, [C.c_char_p, C.POINTER(u64)]),
        "mtl_init": (i32, [u64, u64, u64, C.c_char_p, C.c_size_t]),
        "mtl_buffer_init": (i32, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
        "mtl_budgets": (i32, [C.POINTER(u64), C.POINTER(u64), C.POINTER(u64)]),
        "mtl_shutdown": (i32, []),
        "mtl_last_error": (i32, [C.c_char_p, C.c_size_t]),
        "mtl_device_count": (i32, [C.POINTER(i32)]),
        "mtl_stream_create": (i32, [C.POINTER(vp)]),
        "mtl_stream_destroy": (i32, [vp]),
        "mtl_stream_sync": (i32, [vp]),
        "mtl_stream_native": (vp, [vp]),
        "mtl_alloc": (i32, [u64, C.POINTER(vp)]),
        "mtl_free": (i32, [vp]),
        "mtl_memcpy_h2d": (i32, [vp, vp, u64, vp]),
        "mtl_memcpy_d2h": (i32, [vp, vp, u64, vp]),
        "mtl_filter": (i32, [C.POINTER(Pred), u32, u64, vp, vp, u64, vp, vp, vp]),
        "mtl_filter_noalloc": (i32, [C.POINTER(Pred), u32, u64, vp, vp, u64, vp, vp, vp]),
        "mtl_filter_or": (i32, [C.POINTER(Pred), u32, u32, u64, vp, vp, vp]),
        "mtl_compare": (i32, [C.POINTER(Pred), u32, u64, vp, vp]),
        "mtl_gather": (i32, [vp, u32, vp, vp, u64, vp, vp]),
        "mtl_project": (i32, [vp, i32, vp, vp, u64, i32, i64, vp, vp, vp]),
        "mtl_filter_project": (i32, [C.POINTER(Pred), u32, u64, vp, vp, u64, vp, i32, i32,
                                     i64, vp, vp, vp, vp, vp]),
        "mtl_cache_put": (i32, [u64, u32, vp, u64, u64, C.POINTER(vp)]),
        "mtl_cache_get": (i32, [u64, u32, C.POINTER(vp), C.POINTER(u64)]),
        "mtl_cache_release": (i32, [u64, u32]),
        "mtl_cache_drop": (i32, [u64]),
        "mtl_cache_stats": (i32, [C.POINTER(u64)] * 6),
        "mtl_cache_reset": (i32, []),
        "mtl_catalog_add_table": (i32, [C.c_char_p, u64]),
        "mtl_catalog_add_column": (i32, [C.c_char_p, C.c_char_p, i32, i32, vp, i64, i64, i32]),
        "mtl_catalog_set_dictionary": (i32, [C.c_char_p, C.c_char_p, C.POINTER(C.c_char_p), u32]),
        "mtl_catalog_set_validity": (i32, [C.c_char_p, C.c_char_p, vp]),
        "mtl_catalog_drop": (i32, [C.c_char_p]),
        "mtl_sql_explain": (i32, [C.c_char_p, C.c_char_p, C.c_size_t]),
        "mtl_sql_execute": (i32, [C.c_char_p, C.POINTER(vp), vp]),
        "mtl_result_shape": (i32, [vp, C.POINTER(u32), C.POINTER(u64)]),
        "mtl_result_column": (i32, [vp, u32, C.POINTER(C.c_char_p), C.POINTER(i32),
                                    C.POINTER(i32), C.POINTER(vp)]),
        "mtl_result_validity": (i32, [vp, u32, C.POINTER(vp)]),
        "mtl_result_free": (i32, [vp]),
        "mtl_rewrite_query": (i32, [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "mtl_scalar_hook": (i32, [i32, i32, vp, u64, i64, vp]),
        "mtl_sort": (i32, [vp, i32, vp, vp, u64, i32, vp, C.POINTER(u32), vp]),
        "mtl_join_table_bytes": (u64, [i32, u64, u64]),
        "mtl_join_build": (i32, [i32, vp, i32, vp, vp, u64, i64, u64, vp, u64, vp, vp]),
        "mtl_join_probe": (i32, [i32, vp, u64, i64, u64, vp, i32, vp, vp, u64, vp, vp, vp, vp]),
        "mtl_star_topk": (i32, [C.POINTER(StarDesc), C.POINTER(StarRow), C.POINTER(u32), vp]),
        "mtl_star_topk_async": (i32, [C.POINTER(StarDesc), vp, vp, vp]),
        "mtl_star_status": (i32, [u32]),
        "mtl_profile_enable": (i32, [i32]),
        "mtl_profile_reset": (i32, []),
        "mtl_result_export_arrow": (i32, [vp, vp, vp]),
        "mtl_server_poll": (i32, [vp, u64, vp, C.POINTER(i32)]),
        "mtl_buf_alloc": (i32, [u64, C.POINTER(vp)]),
        "mtl_buf_retain": (i32, [vp]),
        "mtl_buf_release": (i32, [vp]),
        "mtl_buf_device": (vp, [vp]),
        "mtl_buf_bytes": (u64, [vp]),
        "mtl_vec_alloc": (i32, [u64, C.POINTER(vp)]),
        "mtl_vec_retain": (i32, [vp]),
        "mtl_vec_release": (i32, [vp]),
        "mtl_vec_rows": (vp, [vp]),
        "mtl_vec_count_ptr": (vp, [vp]),
        "mtl_vec_capacity": (u64, [vp]),
        "mtl_vec_count": (i32, [vp, vp, C.POINTER(u64)]),
        "mtl_filter_vec": (i32, [C.POINTER(Pred), u32, u64, vp, vp, vp]),
        "mtl_catalog_ingest_arrow": (i32, [C.c_char_p, vp, C.POINTER(vp), u32,
                                           C.POINTER(C.c_char_p), u32, C.POINTER(u64),
                                           C.POINTER(u64), C.c_char_p, C.c_size_t]),
        "mtl_tune_get": (i32, [C.c_char_p, C.POINTER(i32)]),
        "mtl_tune_set": (i32, [C.c_char_p, i32]),
        "mtl_tune_reset": (i32, []),
        "mtl_tune_save": (i32, [C.c_char_p]),
        "mtl_tune_load": (i32, [C.c_char_p]),
        "mtl_tune_describe": (i32, [C.c_char_p, C.c_size_t]),
        "mtl_autotune": (i32, [u32, C.c_char_p, C.c_size_t]),
        "mtl_profile_get": (i32, [u32, C.c_char_p, C.c_size_t, C.POINTER(C.c_double),
                                  C.POINTER(u64)]),
        "mtl_group_agg": (i32, [C.POINTER(GroupDesc), C.POINTER(vp), vp]),
        "mtl_groups_count": (i32, [vp, C.POINTER(u64)]),
        "mtl_groups_fetch": (i32, [vp, vp, C.POINTER(vp), u64, C.POINTER(u64)]),
        "mtl_groups_topk": (i32, [vp, u32, i32, u32, vp, C.POINTER(vp), C.POINTER(u32)]),
        "mtl_groups_release": (i32, [vp]),
        "mtl_server_create": (i32, [u32, u32, u64, C.POINTER(vp)]),
        "mtl_server_submit": (i32, [vp, C.POINTER(Query), i64, C.POINTER(u64)]),
        "mtl_server_submit_sql": (i32, [vp, C.c_char_p, i64, C.POINTER(u64)]),
        "mtl_processing_in_use": (i32, [C.POINTER(u64)]),
        "mtl_server_wait": (i32, [vp, u64, C.POINTER(QueryResult)]),
        "mtl_server_destroy": (i32, [vp]),
        "mtl_now_ns": (i64, []),
        "mtl_query_scratch_bytes": (u64, [C.POINTER(Query)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class MettleError(RuntimeError):
    """Raised for a non-OK status; .code holds the status."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Unsupported(MettleError):
    """MTL_UNSUPPORTED: the caller must use the DuckDB CPU path."""


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().mtl_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status == MTL_OK:
        return
    msg = last_error()
    if status == MTL_UNSUPPORTED:
        raise Unsupported(status, msg)
    raise MettleError(status, msg)
