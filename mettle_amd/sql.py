"""Host-side mirror of the extension's SQL surface over the C ABI:
catalog registration, gpu_processing-style execution, gpu_rewrite_query and
the gpu_*_scalar hooks.  Replaces the reference's plan/ + processing.rs
entry points (parser.rs:20-120, executor/mod.rs:44-138,
processing.rs:103-155) for the admitted subset; MTL_UNSUPPORTED raises
abi.Unsupported, the signal to run DuckDB's CPU path."""
from __future__ import annotations

import ctypes as C

import numpy as np

from mettle_amd import abi

_NP = {abi.INTEGER: np.int32, abi.BIGINT: np.int64, abi.FLOAT: np.float32,
       abi.DOUBLE: np.float64, abi.DATE: np.int32, abi.TIMESTAMP: np.int64,
       abi.DECIMAL: np.int64, abi.DICT8: np.uint8, abi.DICT32: np.int32}


def register_table(name: str, n_rows: int, columns) -> None:
    """columns: list of (name, type, device pointer, min, max, unique[, scale])."""
    L = abi.lib()
    abi.check(L.mtl_catalog_add_table(name.encode(), n_rows))
    for col in columns:
        cname, t, dev, mn, mx, unique = col[:6]
        scale = col[6] if len(col) > 6 else 0
        abi.check(L.mtl_catalog_add_column(name.encode(), cname.encode(), t, scale,
                                           C.c_void_p(dev), int(mn), int(mx), 1 if unique else 0))


def set_dictionary(table: str, column: str, values) -> None:
    arr = (C.c_char_p * len(values))(*[v.encode() for v in values])
    abi.check(abi.lib().mtl_catalog_set_dictionary(table.encode(), column.encode(), arr,
                                                   len(values)))


def drop(table: str | None = None) -> None:
    abi.check(abi.lib().mtl_catalog_drop(None if table is None else table.encode()))


def explain(sql: str) -> str:
    """The plan shape, or raises Unsupported with the reason."""
    buf = C.create_string_buffer(512)
    st = abi.lib().mtl_sql_explain(sql.encode(), buf, 512)
    if st == abi.MTL_UNSUPPORTED:
        raise abi.Unsupported(st, buf.value.decode())
    abi.check(st)
    return buf.value.decode()


def rewrite(sql: str) -> str:
    """gpu_rewrite_query(sql)."""
    L = abi.lib()
    need = C.c_size_t(0)
    buf = C.create_string_buffer(4096)
    st = L.mtl_rewrite_query(sql.encode(), buf, 4096, C.byref(need))
    if st != abi.MTL_OK and need.value > 4096:
        buf = C.create_string_buffer(need.value)
        st = L.mtl_rewrite_query(sql.encode(), buf, need.value, C.byref(need))
    abi.check(st)
    return buf.value.decode()


def execute(sql: str):
    """Run through the GPU path: {column name: (type, scale, numpy array)} in
    select order.  HUGEINT sums come back as Python ints."""
    L = abi.lib()
    h = C.c_void_p()
    abi.check(L.mtl_sql_execute(sql.encode(), C.byref(h), None))
    return take_result(h)


def take_result(h):
    """Convert a mtl_result handle (then freed) to {column name: (type,
    scale, values)}."""
    L = abi.lib()
    h = C.c_void_p(h.value if isinstance(h, C.c_void_p) else h)
    try:
        nc, nr = C.c_uint32(), C.c_uint64()
        abi.check(L.mtl_result_shape(h, C.byref(nc), C.byref(nr)))
        out = {}
        for i in range(nc.value):
            name, t, sc, data = C.c_char_p(), C.c_int32(), C.c_int32(), C.c_void_p()
            abi.check(L.mtl_result_column(h, i, C.byref(name), C.byref(t), C.byref(sc),
                                          C.byref(data)))
            n = nr.value
            if t.value == abi.HUGEINT:
                raw = np.ctypeslib.as_array((C.c_uint64 * (2 * n)).from_address(data.value)) \
                    if n else np.zeros(0, np.uint64)
                vals = [(int(np.int64(raw[2 * r + 1])) << 64) | int(raw[2 * r]) for r in range(n)]
                out[name.value.decode()] = (t.value, sc.value, vals)
            else:
                dt = _NP[t.value]
                arr = np.frombuffer((C.c_char * (n * np.dtype(dt).itemsize)).from_address(
                    data.value), dtype=dt).copy() if n else np.zeros(0, dt)
                out[name.value.decode()] = (t.value, sc.value, arr)
        return out
    finally:
        L.mtl_result_free(h)


def scalar_hook(hook: int, values: np.ndarray, type_code: int, scalar) -> np.ndarray:
    """gpu_{add,mul}_scalar / gpu_{eq,lt,le,gt,ge}_scalar over a host vector."""
    values = np.ascontiguousarray(values)
    out = np.empty(len(values), dtype=values.dtype if hook <= abi.HOOK_MUL else np.uint8)
    abi.check(abi.lib().mtl_scalar_hook(hook, type_code, values.ctypes.data, len(values),
                                        abi.scalar_bits(scalar, type_code), out.ctypes.data))
    return out


def execute_arrow(sql: str):
    """Run through the GPU path and hand the result back as a
    pyarrow.RecordBatch through the Arrow C data interface
    (mtl_result_export_arrow)."""
    import pyarrow as pa
    L = abi.lib()
    h = C.c_void_p()
    abi.check(L.mtl_sql_execute(sql.encode(), C.byref(h), None))
    # ArrowSchema (9 words) and ArrowArray (10 words), owned by pyarrow
    # once imported (it calls their release callbacks)
    schema = C.create_string_buffer(72)
    array = C.create_string_buffer(80)
    try:
        abi.check(L.mtl_result_export_arrow(h, C.addressof(schema), C.addressof(array)))
    finally:
        L.mtl_result_free(h)
    return pa.RecordBatch._import_from_c(C.addressof(array), C.addressof(schema))
