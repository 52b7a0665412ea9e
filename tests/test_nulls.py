"""Ingestion (gpu_register_table, pinned.rs:187-453) and NULLs: Arrow
record batches go through mtl_catalog_ingest_arrow into the HBM column cache
and the SQL catalog, then queries follow DuckDB's NULL semantics.  The
reference pins no column with NULLs (pinned.rs:295-413), so there is no
reference fixture for these results: each expectation is computed here with
explicit validity masks in numpy / pandas (comparisons with NULL are not
true, aggregates skip NULLs and are NULL over no valid value, COUNT(col)
counts valid values, NULL keys form one group and sort last, NULL join keys
match nothing)."""
import numpy as np
import pytest

from mettle_amd import abi
from mettle_amd import sql as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
pa = pytest.importorskip("pyarrow")
pd = pytest.importorskip("pandas")


@pytest.fixture(autouse=True)
def runtime():
    L = abi.lib()
    abi.check(L.mtl_init(2 << 30, 8 << 30, 0, None, 0))
    S.drop()
    yield
    S.drop()
    abi.check(L.mtl_shutdown())


def _col(out, name):
    t, sc, v = out[name]
    v = [int(x) for x in v] if t == abi.HUGEINT else np.asarray(v)
    valid = out[name].valid
    return v, (np.ones(len(v), bool) if valid is None else valid)


def test_ingest_types_and_round_trip():
    """Every supported Arrow type lands with its values, NULLs and
    dictionary; unsupported types are skipped and named."""
    rng = np.random.default_rng(1)
    n = 5000
    i32 = rng.integers(-1000, 1000, n).astype(np.int32)
    m32 = rng.random(n) < 0.2
    i8 = rng.integers(-100, 100, n).astype(np.int8)
    i64 = rng.integers(-10**12, 10**12, n)
    f64 = rng.standard_normal(n)
    mf = rng.random(n) < 0.1
    days = rng.integers(0, 20000, n).astype(np.int32)
    ts = rng.integers(0, 10**15, n)
    dec = rng.integers(-10**9, 10**9, n)
    md = rng.random(n) < 0.3
    words = np.array(["delta", "alpha", "charlie", "bravo", "echo"])
    w = words[rng.integers(0, 5, n)]
    mw = rng.random(n) < 0.15
    import decimal
    t = pa.table({
        "a": pa.array(i32, mask=m32),
        "b": pa.array(i8),
        "c": pa.array(i64, type=pa.int64()),
        "f": pa.array(f64, mask=mf),
        "d": pa.array(days, type=pa.date32()),
        "ts": pa.array(ts, type=pa.timestamp("us")),
        "m": pa.array([decimal.Decimal(int(x)).scaleb(-2) for x in dec], type=pa.decimal128(12, 2),
                      mask=md),
        "s": pa.array(w, mask=mw),
        "e": pa.array(w).dictionary_encode(),
        "flag": pa.array(rng.random(n) < 0.5),
    })
    rows, nbytes, skipped = S.ingest_arrow("T", t)
    assert rows == n and skipped == ["flag"] and nbytes > 0
    out = S.execute("SELECT a, b, c, f, d, ts, m, s, e FROM t")
    v, ok = _col(out, "a")
    np.testing.assert_array_equal(ok, ~m32)
    np.testing.assert_array_equal(v[ok], i32[~m32])
    v, ok = _col(out, "b")
    assert ok.all() and out["b"][0] == abi.INTEGER
    np.testing.assert_array_equal(v, i8.astype(np.int32))
    np.testing.assert_array_equal(_col(out, "c")[0], i64)
    v, ok = _col(out, "f")
    np.testing.assert_array_equal(ok, ~mf)
    np.testing.assert_array_equal(v[ok], f64[~mf])
    np.testing.assert_array_equal(_col(out, "d")[0], days)
    np.testing.assert_array_equal(_col(out, "ts")[0], ts)
    v, ok = _col(out, "m")
    assert out["m"][0] == abi.DECIMAL and out["m"][1] == 2
    np.testing.assert_array_equal(ok, ~md)
    np.testing.assert_array_equal(v[ok], dec[~md])
    # sorted dictionary: codes order like the strings
    srt = sorted(set(words))
    v, ok = _col(out, "s")
    np.testing.assert_array_equal(ok, ~mw)
    assert [srt[c] for c in v[ok]] == list(w[~mw])
    v, ok = _col(out, "e")
    assert ok.all() and [srt[c] for c in v] == list(w)
    # Arrow hand-back carries the NULLs as validity bitmaps
    rb = S.execute_arrow("SELECT a, f, m, s FROM t")
    assert rb.column(0).null_count == int(m32.sum())
    assert rb.column(0).to_pylist() == pa.array(i32, mask=m32).to_pylist()
    assert rb.column(1).to_pylist() == pa.array(f64, mask=mf).to_pylist()
    got_s = rb.column(3).to_pylist()
    assert got_s == pa.array(w, mask=mw).to_pylist()
    assert rb.schema.field(0).nullable


def test_string_comparisons_use_the_sorted_dictionary():
    t = pa.table({"s": pa.array(["pear", "apple", None, "fig", "apple"]),
                  "k": pa.array([1, 2, 3, 4, 5], type=pa.int32())})
    S.ingest_arrow("fruit", t)
    out = S.execute("SELECT k FROM fruit WHERE s = 'apple'")
    assert list(_col(out, "k")[0]) == [2, 5]
    out = S.execute("SELECT k FROM fruit WHERE s IS NULL")
    assert list(_col(out, "k")[0]) == [3]
    out = S.execute("SELECT k, s FROM fruit WHERE s IS NOT NULL ORDER BY s")
    assert list(_col(out, "k")[0]) in ([2, 5, 4, 1], [5, 2, 4, 1])
    # range comparisons through the sorted dictionary
    assert list(_col(S.execute("SELECT k FROM fruit WHERE s < 'fig'"), "k")[0]) == [2, 5]
    assert list(_col(S.execute("SELECT k FROM fruit WHERE s <= 'fig'"), "k")[0]) == [2, 4, 5]
    assert list(_col(S.execute("SELECT k FROM fruit WHERE s > 'banana'"), "k")[0]) == [1, 4]
    assert list(_col(S.execute("SELECT k FROM fruit WHERE s >= 'pear'"), "k")[0]) == [1]
    assert list(_col(S.execute("SELECT k FROM fruit WHERE s = 'kiwi'"), "k")[0]) == []
    out = S.execute("SELECT k FROM fruit WHERE s BETWEEN 'b' AND 'g'")
    assert list(_col(out, "k")[0]) == [4]


@pytest.mark.parametrize("p_null", [0.0, 0.3, 1.0])
def test_filters_with_nulls(p_null):
    rng = np.random.default_rng(2)
    n = 200_000
    x = rng.integers(-100, 100, n).astype(np.int64)
    y = rng.integers(0, 10, n).astype(np.int32)
    mx = rng.random(n) < p_null
    S.ingest_arrow("t", pa.table({"x": pa.array(x, mask=mx), "y": pa.array(y)}))
    vx = ~mx
    cases = [
        ("x > 5", vx & (x > 5)),
        ("x IS NULL", mx),
        ("x IS NOT NULL AND y = 3", vx & (y == 3)),
        ("x IN (1, 2, 3)", vx & np.isin(x, [1, 2, 3])),
        ("(x < -90 OR x > 90) AND y < 5", vx & ((x < -90) | (x > 90)) & (y < 5)),
        ("y = 1 OR y = 2", (y == 1) | (y == 2)),
    ]
    for where, m in cases:
        out = S.execute(f"SELECT y FROM t WHERE {where}")
        assert len(out["y"][2]) == int(m.sum()), where
        np.testing.assert_array_equal(out["y"][2], y[m])
    if p_null > 0:  # OR over two columns, one with NULLs
        with pytest.raises(abi.Unsupported):
            S.execute("SELECT y FROM t WHERE x = 1 OR y = 2")


def test_order_by_puts_nulls_last():
    rng = np.random.default_rng(3)
    n = 50_000
    x = rng.integers(-1000, 1000, n).astype(np.int32)
    mx = rng.random(n) < 0.25
    S.ingest_arrow("t", pa.table({"x": pa.array(x, mask=mx),
                                  "r": pa.array(np.arange(n, dtype=np.int32))}))
    for desc in (False, True):
        out = S.execute("SELECT r, x FROM t WHERE r >= 10 ORDER BY x" + (" DESC" if desc else ""))
        r = np.asarray(out["r"][2])
        nv = int((~mx[10:]).sum())
        xs = x[r[:nv]]
        assert (~mx[r[:nv]]).all() and mx[r[nv:]].all()
        assert (np.diff(xs.astype(np.int64)) <= 0).all() if desc else (np.diff(xs) >= 0).all()
        assert list(r[nv:]) == list(np.flatnonzero(mx[10:]) + 10)  # NULLs in row order
        assert not out["x"].valid[nv:].any()


def _group_expect(keys, kvalid, vals, vvalid):
    """Per group key (None = NULL group): count(*), count(v), sum, min, max
    over the valid values."""
    g = {}
    for k, kv, v, vv in zip(keys, kvalid, vals, vvalid):
        kk = int(k) if kv else None
        e = g.setdefault(kk, [0, 0, 0, None, None])
        e[0] += 1
        if vv:
            e[1] += 1
            e[2] += int(v)
            e[3] = int(v) if e[3] is None else min(e[3], int(v))
            e[4] = int(v) if e[4] is None else max(e[4], int(v))
    return g


def test_group_by_with_null_keys_and_values():
    rng = np.random.default_rng(4)
    n = 100_000
    k = rng.integers(0, 50, n).astype(np.int32)
    mk = rng.random(n) < 0.1
    v = rng.integers(-10**6, 10**6, n)
    mv = rng.random(n) < 0.3
    mv[k == 7] = True  # a group whose values are all NULL
    S.ingest_arrow("t", pa.table({"k": pa.array(k, mask=mk), "v": pa.array(v, mask=mv)}))
    out = S.execute("SELECT k, COUNT(*) AS n, COUNT(v) AS nv, SUM(v) AS s, MIN(v) AS lo, "
                    "MAX(v) AS hi, AVG(v) AS av FROM t GROUP BY k ORDER BY k")
    exp = _group_expect(k, ~mk, v, ~mv)
    order = sorted(x for x in exp if x is not None) + [None]  # NULL key last
    kk, kok = _col(out, "k")
    assert [int(a) if ok else None for a, ok in zip(kk, kok)] == order
    assert list(_col(out, "n")[0]) == [exp[g][0] for g in order]
    assert list(_col(out, "nv")[0]) == [exp[g][1] for g in order]
    s, sok = _col(out, "s")
    assert [a if ok else None for a, ok in zip(s, sok)] == \
        [exp[g][2] if exp[g][1] else None for g in order]
    for name, idx in (("lo", 3), ("hi", 4)):
        a, ok = _col(out, name)
        assert [int(x) if o else None for x, o in zip(a, ok)] == [exp[g][idx] for g in order]
    av, aok = _col(out, "av")
    for g, x, ok in zip(order, av, aok):
        if exp[g][1] == 0:
            assert not ok
        else:
            assert ok and x == pytest.approx(exp[g][2] / exp[g][1], rel=1e-15)


def test_float_aggregates_skip_nulls():
    rng = np.random.default_rng(5)
    n = 40_000
    k = rng.integers(0, 8, n).astype(np.int32)
    f = rng.standard_normal(n)
    f[rng.random(n) < 0.01] = np.nan
    mf = rng.random(n) < 0.4
    S.ingest_arrow("t", pa.table({"k": pa.array(k), "f": pa.array(f, mask=mf)}))
    out = S.execute("SELECT k, MIN(f) AS lo, MAX(f) AS hi, SUM(f) AS s, COUNT(f) AS c "
                    "FROM t GROUP BY k ORDER BY k")
    for i, g in enumerate(range(8)):
        sel = (k == g) & ~mf
        fv = f[sel]
        lo = np.nan if np.isnan(fv).all() else np.nanmin(fv)
        hi = np.nan if np.isnan(fv).any() else fv.max()  # NaN is the largest value
        assert _col(out, "c")[0][i] == sel.sum()
        got_lo = _col(out, "lo")[0][i]
        got_hi = _col(out, "hi")[0][i]
        assert (np.isnan(got_lo) and np.isnan(lo)) or got_lo == lo
        assert (np.isnan(got_hi) and np.isnan(hi)) or got_hi == hi
        s = _col(out, "s")[0][i]
        assert (np.isnan(s) and np.isnan(fv).any()) or s == pytest.approx(fv.sum(), abs=1e-9)


def test_joins_skip_null_keys():
    rng = np.random.default_rng(6)
    na, nb = 3000, 80_000
    ak = np.arange(na, dtype=np.int32)
    mak = rng.random(na) < 0.1
    ag = rng.integers(0, 5, na).astype(np.int32)
    bk = rng.integers(0, na, nb).astype(np.int32)
    mbk = rng.random(nb) < 0.2
    bv = rng.integers(0, 1000, nb)
    mbv = rng.random(nb) < 0.1
    S.ingest_arrow("a", pa.table({"k": pa.array(ak, mask=mak), "g": pa.array(ag)}), unique=["k"])
    S.ingest_arrow("b", pa.table({"k": pa.array(bk, mask=mbk), "v": pa.array(bv, mask=mbv)}))
    out = S.execute("SELECT a.g, COUNT(*) AS n, SUM(b.v) AS s, COUNT(b.v) AS c FROM a "
                    "JOIN b ON a.k = b.k GROUP BY a.g ORDER BY a.g")
    ok_b = ~mbk & ~mak[bk]  # both keys valid
    for i, g in enumerate(range(5)):
        sel = ok_b & (ag[bk] == g)
        assert _col(out, "n")[0][i] == sel.sum()
        assert _col(out, "s")[0][i] == int(bv[sel & ~mbv].sum())
        assert _col(out, "c")[0][i] == int((sel & ~mbv).sum())
    # projection: NULL values come back as NULLs
    out = S.execute("SELECT b.v, a.g FROM a JOIN b ON a.k = b.k WHERE a.g = 2 ORDER BY b.v")
    sel = ok_b & (ag[bk] == 2)
    v, ok = _col(out, "v")
    nv = int((sel & ~mbv).sum())
    assert ok[:nv].all() and not ok[nv:].any()
    np.testing.assert_array_equal(v[:nv], np.sort(bv[sel & ~mbv]))


def test_ungrouped_aggregates_over_no_rows_are_null():
    S.ingest_arrow("t", pa.table({"x": pa.array([1, 2, None], type=pa.int64())}))
    out = S.execute("SELECT COUNT(*) AS n, COUNT(x) AS c, SUM(x) AS s, MIN(x) AS lo, "
                    "AVG(x) AS av FROM t WHERE x > 5")
    assert list(_col(out, "n")[0]) == [0] and list(_col(out, "c")[0]) == [0]
    for name in ("s", "lo", "av"):
        assert not _col(out, name)[1][0]
    out = S.execute("SELECT COUNT(*) AS n, COUNT(x) AS c, SUM(x) AS s FROM t")
    assert [_col(out, c)[0][0] for c in ("n", "c", "s")] == [3, 2, 3]


def test_ingested_tables_run_the_star_query():
    """Columns without NULLs through ingestion keep the star pipeline
    (declared unique keys), and the result matches the generic joins."""
    rng = np.random.default_rng(7)
    nc, no, nl = 2000, 20_000, 100_000
    seg = np.array(["AUTOMOBILE", "BUILDING", "FURNITURE", "MACHINERY", "HOUSEHOLD"])
    cust = pa.table({"c_key": pa.array(np.arange(1, nc + 1, dtype=np.int32)),
                     "segment": pa.array(seg[rng.integers(0, 5, nc)])})
    okey = np.arange(no, dtype=np.int32) * 4 + 1
    orders = pa.table({"o_key": pa.array(okey),
                       "o_cust": pa.array(rng.integers(1, nc + 1, no).astype(np.int32)),
                       "o_date": pa.array(rng.integers(8000, 10000, no).astype(np.int32),
                                          type=pa.date32()),
                       "o_prio": pa.array(rng.integers(0, 5, no).astype(np.int32))})
    lines = pa.table({"l_order": pa.array(okey[rng.integers(0, no, nl)]),
                      "price": pa.array(rng.integers(100, 10**6, nl)),
                      "discount": pa.array(rng.integers(0, 11, nl)),
                      "l_ship": pa.array(rng.integers(8000, 10000, nl).astype(np.int32),
                                         type=pa.date32())})
    S.ingest_arrow("customers", cust, unique=["c_key"])
    S.ingest_arrow("orders", orders, unique=["o_key"])
    S.ingest_arrow("lines", lines)
    q = ("SELECT l.l_order, SUM(l.price * (100 - l.discount)) AS revenue, o.o_date, o.o_prio "
         "FROM customers c, orders o, lines l WHERE c.segment = 'BUILDING' AND "
         "c.c_key = o.o_cust AND l.l_order = o.o_key AND o.o_date < DATE '1995-03-15' AND "
         "l.l_ship > DATE '1995-03-15' GROUP BY l.l_order, o.o_date, o.o_prio "
         "ORDER BY revenue DESC, o.o_date LIMIT 10")
    assert S.explain(q) in ("star2", "star3")
    star = S.execute(q)
    q2 = q.replace("c.segment = 'BUILDING'", "c.segment IN ('BUILDING', 'NO SUCH')")
    assert S.explain(q2) == "join"
    gen = S.execute(q2)
    assert [int(x) for x in star["l_order"][2]] == [int(x) for x in gen["l_order"][2]]
    assert [int(x) for x in star["revenue"][2]] == [int(x) for x in gen["revenue"][2]]
