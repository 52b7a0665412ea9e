"""Arrow hand-back: results exported through the Arrow C data interface
(mtl_result_export_arrow) and imported by pyarrow carry the result's values
and DuckDB's Arrow types (decimal128 for DECIMAL and HUGEINT sums,
dictionary<utf8> for dictionary columns), equal to the ctypes view of the
same result."""
import decimal

import numpy as np
import pytest

from mettle_amd import abi
from mettle_amd import sql as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
pa = pytest.importorskip("pyarrow")


@pytest.fixture(scope="module")
def ops():
    from mettle_amd import ops as _ops
    return _ops


@pytest.fixture(autouse=True)
def clean_catalog():
    S.drop()
    yield
    S.drop()


def _register(ops, n=5000, seed=3):
    rng = np.random.default_rng(seed)
    cols = {
        "i": (abi.INTEGER, rng.integers(-2**31, 2**31, n).astype(np.int32)),
        "b": (abi.BIGINT, rng.integers(-2**62, 2**62, n).astype(np.int64)),
        "f": (abi.FLOAT, rng.standard_normal(n).astype(np.float32)),
        "g": (abi.DOUBLE, rng.standard_normal(n)),
        "d": (abi.DATE, rng.integers(10957, 22281, n).astype(np.int32)),
        "t": (abi.TIMESTAMP, rng.integers(946684800 * 10**6, 1924991999 * 10**6, n)
              .astype(np.int64)),
        "p": (abi.DECIMAL, rng.integers(-10**9, 10**9, n).astype(np.int64)),
        "s": (abi.DICT8, rng.integers(0, 3, n).astype(np.uint8)),
        "k": (abi.INTEGER, rng.integers(0, 50, n).astype(np.int32)),
    }
    keep, spec = [], []
    for name, (t, v) in cols.items():
        d = ops.to_device(v)
        keep.append(d)
        mn, mx = (0, 0) if t in (abi.FLOAT, abi.DOUBLE) else (int(v.min()), int(v.max()))
        spec.append((name, t, d.data_ptr(), mn, mx, False) + ((2,) if t == abi.DECIMAL else ()))
    S.register_table("x", n, spec)
    S.set_dictionary("x", "s", ["red", "green", "blue"])
    return cols, keep


def test_scan_result_as_arrow(ops):
    cols, keep = _register(ops)
    q = "SELECT i, b, f, g, d, t, p, s FROM x WHERE k < 20 ORDER BY b"
    rb = S.execute_arrow(q)
    ref = S.execute(q)
    assert rb.schema.names == list(ref)
    assert rb.num_rows == len(ref["i"][2])
    types = {f.name: f.type for f in rb.schema}
    assert types["i"] == pa.int32() and types["b"] == pa.int64()
    assert types["f"] == pa.float32() and types["g"] == pa.float64()
    assert types["d"] == pa.date32() and types["t"] == pa.timestamp("us")
    assert types["p"] == pa.decimal128(18, 2)
    assert types["s"] == pa.dictionary(pa.uint8(), pa.utf8())
    for name in ("i", "b", "f", "g"):
        np.testing.assert_array_equal(rb.column(name).to_numpy(), ref[name][2])
    np.testing.assert_array_equal(rb.column("d").cast(pa.int32()).to_numpy(), ref["d"][2])
    np.testing.assert_array_equal(rb.column("t").cast(pa.int64()).to_numpy(), ref["t"][2])
    assert [int(x.scaleb(2)) for x in rb.column("p").to_pylist()] == [int(v) for v in ref["p"][2]]
    names = ["red", "green", "blue"]
    assert rb.column("s").to_pylist() == [names[c] for c in ref["s"][2]]


def test_group_result_as_arrow(ops):
    cols, keep = _register(ops, n=20000, seed=4)
    q = ("SELECT k, SUM(b) AS sb, SUM(p) AS sp, COUNT(*) AS n, AVG(g) AS ag, MIN(d) AS md "
         "FROM x GROUP BY k ORDER BY k")
    rb = S.execute_arrow(q)
    types = {f.name: f.type for f in rb.schema}
    assert types["sb"] == pa.decimal128(38, 0) and types["sp"] == pa.decimal128(38, 2)
    assert types["n"] == pa.int64() and types["ag"] == pa.float64()
    k, b, p = cols["k"][1], cols["b"][1], cols["p"][1]
    keys = sorted(set(int(x) for x in k))
    assert rb.column("k").to_pylist() == keys
    assert [int(x) for x in rb.column("sb").to_pylist()] == \
        [int(sum(int(v) for v in b[k == g])) for g in keys]
    assert [x for x in rb.column("sp").to_pylist()] == \
        [decimal.Decimal(int(p[k == g].sum())).scaleb(-2) for g in keys]
    assert rb.column("n").to_pylist() == [int((k == g).sum()) for g in keys]


def test_exported_batch_outlives_the_result(ops):
    cols, keep = _register(ops, n=1000, seed=5)
    rb = S.execute_arrow("SELECT i, s FROM x WHERE k = 7")
    import gc
    gc.collect()
    sel = np.flatnonzero(cols["k"][1] == 7)
    np.testing.assert_array_equal(rb.column("i").to_numpy(), cols["i"][1][sel])
    assert rb.column("s").to_pylist() == [["red", "green", "blue"][c] for c in cols["s"][1][sel]]
