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


def ingest_arrow(name: str, data, unique=()):
    """gpu_register_table over a pyarrow Table or RecordBatch (or a list of
    batches): mtl_catalog_ingest_arrow through the Arrow C data interface.
    Returns (rows, bytes, skipped column names).  Needs gpu_buffer_init."""
    import pyarrow as pa
    batches = data.to_batches() if isinstance(data, pa.Table) else (
        [data] if isinstance(data, pa.RecordBatch) else list(data))
    schema = data.schema if not isinstance(data, list) else batches[0].schema
    s_buf = C.create_string_buffer(72)
    a_bufs = [C.create_string_buffer(80) for _ in batches]
    schema._export_to_c(C.addressof(s_buf))
    for b, ab in zip(batches, a_bufs):
        b._export_to_c(C.addressof(ab))
    try:
        arr = (C.c_void_p * max(1, len(batches)))(*[C.addressof(ab) for ab in a_bufs])
        uq = (C.c_char_p * max(1, len(unique)))(*[u.encode() for u in unique])
        rows, nbytes = C.c_uint64(), C.c_uint64()
        skipped = C.create_string_buffer(4096)
        abi.check(abi.lib().mtl_catalog_ingest_arrow(
            name.encode(), C.addressof(s_buf), arr, len(batches), uq, len(unique),
            C.byref(rows), C.byref(nbytes), skipped, 4096))
    finally:
        # hand the exported structs back to pyarrow, which releases them
        for ab in a_bufs:
            pa.RecordBatch._import_from_c(C.addressof(ab), schema)
        pa.Schema._import_from_c(C.addressof(s_buf))
    sk = skipped.value.decode()
    return rows.value, nbytes.value, [x for x in sk.split(",") if x]


def set_validity(table: str, column: str, valid_dev_ptr) -> None:
    """NULLs of a registered column: device bytes, one per row, 1 = valid
    (0 / None: no NULLs).  Value bytes of NULL rows must be zero."""
    abi.check(abi.lib().mtl_catalog_set_validity(table.encode(), column.encode(),
                                                 C.c_void_p(valid_dev_ptr or None)))


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


class Column(tuple):
    """(type, scale, values) of a result column; .valid is a numpy bool
    array (True = valid) when the column holds NULLs, else None."""
    valid = None


def execute(sql: str):
    """Run through the GPU path: {column name: Column(type, scale, values)}
    in select order.  HUGEINT sums come back as Python ints; NULL rows are
    marked in each Column's .valid."""
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
                col = Column((t.value, sc.value, vals))
            else:
                dt = _NP[t.value]
                arr = np.frombuffer((C.c_char * (n * np.dtype(dt).itemsize)).from_address(
                    data.value), dtype=dt).copy() if n else np.zeros(0, dt)
                col = Column((t.value, sc.value, arr))
            vp = C.c_void_p()
            abi.check(L.mtl_result_validity(h, i, C.byref(vp)))
            if vp.value and n:
                col.valid = np.frombuffer((C.c_char * n).from_address(vp.value),
                                          dtype=np.uint8).astype(bool)
            out[name.value.decode()] = col
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
