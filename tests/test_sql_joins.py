"""The generic join lowering: the reference's SQL fixtures for ungrouped
aggregates, a two-table join with a projection, and joins grouped by
non-key columns (tests/extension.rs:239-276, 279-317, 475-525, 973-1016,
expected values as written there), then larger random joins checked
against numpy/pandas (merge, groupby) on the host, and many-to-many joins
(keys repeating on both sides)."""
import numpy as np
import pytest

from mettle_amd import abi
from mettle_amd import sql as S

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
pd = pytest.importorskip("pandas")


@pytest.fixture(scope="module")
def ops():
    from mettle_amd import ops as _ops
    return _ops


@pytest.fixture(autouse=True)
def clean_catalog():
    S.drop()
    yield
    S.drop()


_keep = []


def _table(ops, name, cols):
    """cols: list of (name, type, numpy values[, unique]); registers the
    table with each column's [min, max]."""
    n = len(cols[0][2])
    spec = []
    for c in cols:
        cname, t, v = c[:3]
        unique = c[3] if len(c) > 3 else False
        scale = c[4] if len(c) > 4 else 0
        d = ops.to_device(np.ascontiguousarray(v))
        _keep.append(d)
        mn, mx = (int(v.min()), int(v.max())) if n and t not in (abi.FLOAT, abi.DOUBLE) else (0, 0)
        spec.append((cname, t, d.data_ptr(), mn, mx, unique, scale))
    S.register_table(name, n, spec)


def _ints(out, name):
    t, sc, v = out[name]
    if t == abi.HUGEINT:
        return [int(x) for x in v]
    return [int(x) for x in np.asarray(v)]


def test_reference_ungrouped_aggregate(ops):
    # tests/extension.rs:239-276
    _table(ops, "prices", [("value", abi.INTEGER, np.array([10, 20, 30, 40, 50], np.int32))])
    out = S.execute("SELECT SUM(value) AS total, COUNT(*) AS cnt, MIN(value) AS lo, "
                    "MAX(value) AS hi FROM prices WHERE value >= 20")
    assert list(out) == ["total", "cnt", "lo", "hi"]
    assert (_ints(out, "total"), _ints(out, "cnt"), _ints(out, "lo"), _ints(out, "hi")) == \
        ([140], [4], [20], [50])


def test_reference_two_table_join(ops):
    # tests/extension.rs:279-317 (dept_name VARCHAR, dictionary-coded here)
    _table(ops, "departments", [("dept_id", abi.INTEGER, np.array([1, 2, 3], np.int32), True),
                                ("dept_name", abi.DICT8, np.array([0, 1, 2], np.uint8))])
    S.set_dictionary("departments", "dept_name", ["Engineering", "Sales", "HR"])
    _table(ops, "employees", [("emp_id", abi.INTEGER, np.array([10, 11, 12, 13], np.int32), True),
                              ("dept_id", abi.INTEGER, np.array([1, 1, 2, 3], np.int32)),
                              ("salary", abi.INTEGER, np.array([100, 200, 150, 300], np.int32))])
    out = S.execute("SELECT e.emp_id, e.salary FROM departments d JOIN employees e "
                    "ON d.dept_id = e.dept_id WHERE d.dept_name = 'Engineering'")
    rows = sorted(zip(_ints(out, "emp_id"), _ints(out, "salary")))
    assert rows == [(10, 100), (11, 200)]


def test_reference_two_table_grouped_aggregate(ops):
    # tests/extension.rs:973-1016
    _table(ops, "departments", [("id", abi.INTEGER, np.array([1, 2, 3], np.int32), True),
                                ("region", abi.INTEGER, np.array([10, 10, 20], np.int32))])
    _table(ops, "employees", [("dept_id", abi.INTEGER, np.array([1, 1, 2, 3, 3], np.int32)),
                              ("salary", abi.DOUBLE,
                               np.array([100.0, 150.0, 200.0, 90.0, 110.0]))])
    out = S.execute("SELECT d.region, SUM(e.salary) AS payroll FROM departments d "
                    "JOIN employees e ON d.id = e.dept_id GROUP BY d.region ORDER BY d.region")
    assert _ints(out, "region") == [10, 20]
    assert list(out["payroll"][2]) == [450.0, 200.0]


def test_reference_three_table_grouped_aggregate(ops):
    # tests/extension.rs:475-525
    _table(ops, "regions", [("region_id", abi.INTEGER, np.array([10, 20], np.int32), True),
                            ("market", abi.INTEGER, np.array([1, 2], np.int32))])
    _table(ops, "departments", [("dept_id", abi.INTEGER, np.array([1, 2, 3], np.int32), True),
                                ("region_id", abi.INTEGER, np.array([10, 10, 20], np.int32))])
    _table(ops, "employees", [("emp_id", abi.INTEGER, np.arange(100, 105, dtype=np.int32), True),
                              ("dept_id", abi.INTEGER, np.array([1, 1, 2, 3, 3], np.int32)),
                              ("salary", abi.DOUBLE,
                               np.array([100.0, 150.0, 200.0, 90.0, 110.0]))])
    out = S.execute("SELECT r.market, SUM(e.salary) AS payroll FROM regions r "
                    "JOIN departments d ON r.region_id = d.region_id "
                    "JOIN employees e ON d.dept_id = e.dept_id "
                    "GROUP BY r.market ORDER BY r.market")
    assert _ints(out, "market") == [1, 2]
    assert list(out["payroll"][2]) == [450.0, 200.0]


def _star(ops, seed, n_c=3000, n_o=40_000, n_l=200_000):
    rng = np.random.default_rng(seed)
    cust = {"c_key": rng.permutation(n_c).astype(np.int32) + 1,
            "c_region": rng.integers(0, 7, n_c).astype(np.int32)}
    orders = {"o_key": rng.permutation(n_o).astype(np.int32) * 3 + 5,
              "o_cust": rng.integers(1, n_c + 1, n_o).astype(np.int32),
              "o_prio": rng.integers(0, 5, n_o).astype(np.int32)}
    lines = {"l_order": orders["o_key"][rng.integers(0, n_o, n_l)],
             "l_qty": rng.integers(-50, 51, n_l).astype(np.int64),
             "l_disc": rng.integers(0, 11, n_l).astype(np.int32)}
    _table(ops, "cust", [("c_key", abi.INTEGER, cust["c_key"], True),
                         ("c_region", abi.INTEGER, cust["c_region"])])
    _table(ops, "orders", [("o_key", abi.INTEGER, orders["o_key"], True),
                           ("o_cust", abi.INTEGER, orders["o_cust"]),
                           ("o_prio", abi.INTEGER, orders["o_prio"])])
    _table(ops, "lines", [("l_order", abi.INTEGER, lines["l_order"]),
                          ("l_qty", abi.BIGINT, lines["l_qty"]),
                          ("l_disc", abi.INTEGER, lines["l_disc"])])
    return pd.DataFrame(cust), pd.DataFrame(orders), pd.DataFrame(lines)


def test_three_table_join_grouped_by_non_key_columns(ops):
    c, o, l = _star(ops, 1)
    out = S.execute("SELECT c.c_region, o.o_prio, SUM(l.l_qty) AS q, COUNT(*) AS n, "
                    "MIN(l.l_disc) AS lo, MAX(l.l_qty) AS hi "
                    "FROM lines l JOIN orders o ON l.l_order = o.o_key "
                    "JOIN cust c ON o.o_cust = c.c_key "
                    "WHERE o.o_prio < 4 AND l.l_disc >= 2 "
                    "GROUP BY c.c_region, o.o_prio ORDER BY c.c_region, o.o_prio")
    j = l[l.l_disc >= 2].merge(o[o.o_prio < 4], left_on="l_order", right_on="o_key") \
        .merge(c, left_on="o_cust", right_on="c_key")
    g = j.groupby(["c_region", "o_prio"]).agg(q=("l_qty", "sum"), n=("l_qty", "size"),
                                              lo=("l_disc", "min"), hi=("l_qty", "max"))
    g = g.reset_index().sort_values(["c_region", "o_prio"])
    assert _ints(out, "c_region") == list(g.c_region)
    assert _ints(out, "o_prio") == list(g.o_prio)
    assert _ints(out, "q") == [int(x) for x in g.q]
    assert _ints(out, "n") == [int(x) for x in g.n]
    assert _ints(out, "lo") == [int(x) for x in g.lo]
    assert _ints(out, "hi") == [int(x) for x in g.hi]
    # descending key order: the same groups reversed
    out_d = S.execute("SELECT c.c_region, SUM(l.l_qty) AS q FROM lines l "
                      "JOIN orders o ON l.l_order = o.o_key JOIN cust c ON o.o_cust = c.c_key "
                      "GROUP BY c.c_region ORDER BY c.c_region DESC")
    g2 = l.merge(o, left_on="l_order", right_on="o_key").merge(c, left_on="o_cust",
                                                               right_on="c_key")
    g2 = g2.groupby("c_region").l_qty.sum().sort_index(ascending=False)
    assert _ints(out_d, "c_region") == list(g2.index)
    assert _ints(out_d, "q") == [int(x) for x in g2]


def test_join_projection_with_order_and_limit(ops):
    c, o, l = _star(ops, 2)
    out = S.execute("SELECT o.o_key, c.c_region, o.o_prio FROM orders o "
                    "JOIN cust c ON o.o_cust = c.c_key WHERE c.c_region = 3 AND o.o_prio >= 2 "
                    "ORDER BY o.o_key DESC LIMIT 500")
    j = o[o.o_prio >= 2].merge(c[c.c_region == 3], left_on="o_cust", right_on="c_key")
    j = j.sort_values("o_key", ascending=False, kind="stable").head(500)
    assert _ints(out, "o_key") == list(j.o_key)
    assert _ints(out, "c_region") == list(j.c_region)
    assert _ints(out, "o_prio") == list(j.o_prio)


def test_ungrouped_aggregates_over_a_join(ops):
    c, o, l = _star(ops, 3)
    out = S.execute("SELECT COUNT(*) AS n, SUM(l.l_qty) AS q, MAX(o.o_prio) AS p "
                    "FROM lines l JOIN orders o ON l.l_order = o.o_key WHERE o.o_prio = 1")
    j = l.merge(o[o.o_prio == 1], left_on="l_order", right_on="o_key")
    assert (_ints(out, "n"), _ints(out, "q"), _ints(out, "p")) == \
        ([len(j)], [int(j.l_qty.sum())], [1])


def test_many_to_many_join_small(ops):
    """Keys repeating on both sides: every matching pair (DuckDB's inner
    join), not the reference's overwritten build keys (kernels.metal:202-206)."""
    _table(ops, "a", [("k", abi.INTEGER, np.array([1, 1, 2, 3], np.int32)),
                      ("v", abi.INTEGER, np.array([5, 6, 7, 4], np.int32))])
    _table(ops, "b", [("k", abi.INTEGER, np.array([1, 2, 2, 1, 9], np.int32)),
                      ("w", abi.INTEGER, np.array([8, 9, 10, 11, 12], np.int32))])
    out = S.execute("SELECT a.v, b.w FROM a JOIN b ON a.k = b.k")
    got = sorted(zip(_ints(out, "v"), _ints(out, "w")))
    assert got == [(5, 8), (5, 11), (6, 8), (6, 11), (7, 9), (7, 10)]
    out = S.execute("SELECT a.k, COUNT(*) AS n, SUM(b.w) AS s FROM a JOIN b ON a.k = b.k "
                    "GROUP BY a.k ORDER BY a.k")
    assert _ints(out, "k") == [1, 2]
    assert _ints(out, "n") == [4, 2]
    assert _ints(out, "s") == [38, 19]


@pytest.mark.parametrize("seed,na,nb,dom", [(1, 50_000, 60_000, 20_000),
                                            (2, 200_000, 3_000, 1_000),
                                            (3, 5_000, 400_000, 70)])
def test_many_to_many_join_random(ops, seed, na, nb, dom):
    """Random many-to-many joins with heavy duplication on both sides,
    filtered, projected and grouped, against pandas merge."""
    rng = np.random.default_rng(seed)
    ak = (rng.integers(0, dom, na) * 7 - 1000).astype(np.int64)
    av = rng.integers(-10**6, 10**6, na).astype(np.int64)
    bk = (rng.integers(0, dom, nb) * 7 - 1000).astype(np.int64)
    bw = rng.integers(0, 100, nb).astype(np.int32)
    _table(ops, "a", [("k", abi.BIGINT, ak), ("v", abi.BIGINT, av)])
    _table(ops, "b", [("k", abi.BIGINT, bk), ("w", abi.INTEGER, bw)])
    A = pd.DataFrame({"k": ak, "v": av})
    B = pd.DataFrame({"k": bk, "w": bw})
    j = A[A.v > 0].merge(B[B.w < 30], on="k")
    out = S.execute("SELECT b.w, COUNT(*) AS n, SUM(a.v) AS s, MIN(a.v) AS lo FROM a "
                    "JOIN b ON a.k = b.k WHERE a.v > 0 AND b.w < 30 GROUP BY b.w ORDER BY b.w")
    g = j.groupby("w").v.agg(["size", "sum", "min"])
    assert _ints(out, "w") == list(g.index)
    assert _ints(out, "n") == [int(x) for x in g["size"]]
    assert _ints(out, "s") == [int(x) for x in g["sum"]]
    assert _ints(out, "lo") == [int(x) for x in g["min"]]
    if len(j) < 3_000_000:
        out = S.execute("SELECT a.v, b.w FROM a JOIN b ON a.k = b.k WHERE a.v > 0 AND b.w < 30")
        got = np.stack([np.asarray(_ints(out, "v")), np.asarray(_ints(out, "w"))], 1)
        exp = j[["v", "w"]].to_numpy()
        assert len(got) == len(exp)
        got = got[np.lexsort(got.T[::-1])]
        exp = exp[np.lexsort(exp.T[::-1])]
        np.testing.assert_array_equal(got, exp)


def test_many_to_many_three_tables(ops):
    """A chain a-b-c where both joins repeat keys on both sides."""
    rng = np.random.default_rng(21)
    a = pd.DataFrame({"x": rng.integers(0, 300, 4000).astype(np.int32),
                      "v": rng.integers(0, 1000, 4000).astype(np.int64)})
    b = pd.DataFrame({"x": rng.integers(0, 300, 3000).astype(np.int32),
                      "y": rng.integers(0, 50, 3000).astype(np.int32)})
    c = pd.DataFrame({"y": rng.integers(0, 50, 2000).astype(np.int32),
                      "g": rng.integers(0, 9, 2000).astype(np.int32)})
    _table(ops, "a", [("x", abi.INTEGER, a.x.to_numpy()), ("v", abi.BIGINT, a.v.to_numpy())])
    _table(ops, "b", [("x", abi.INTEGER, b.x.to_numpy()), ("y", abi.INTEGER, b.y.to_numpy())])
    _table(ops, "c", [("y", abi.INTEGER, c.y.to_numpy()), ("g", abi.INTEGER, c.g.to_numpy())])
    out = S.execute("SELECT c.g, COUNT(*) AS n, SUM(a.v) AS s FROM a JOIN b ON a.x = b.x "
                    "JOIN c ON b.y = c.y WHERE a.v < 500 GROUP BY c.g ORDER BY c.g")
    j = a[a.v < 500].merge(b, on="x").merge(c, on="y")
    g = j.groupby("g").v.agg(["size", "sum"])
    assert _ints(out, "g") == list(g.index)
    assert _ints(out, "n") == [int(x) for x in g["size"]]
    assert _ints(out, "s") == [int(x) for x in g["sum"]]


def test_or_and_in_predicates(ops):
    """IN lists and parenthesised ORs (a conjunction of disjunctions) in a
    scan, a group-by and a join, against numpy."""
    rng = np.random.default_rng(11)
    n = 300_000
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    w = rng.integers(0, 50, n).astype(np.int64)
    d = rng.integers(10957, 22281, n).astype(np.int32)
    _table(ops, "t", [("v", abi.INTEGER, v), ("w", abi.BIGINT, w), ("d", abi.DATE, d)])
    out = S.execute("SELECT v, w FROM t WHERE w IN (3, 7, 11) AND (v < -900 OR v >= 950) "
                    "ORDER BY v")
    m = np.isin(w, [3, 7, 11]) & ((v < -900) | (v >= 950))
    sel = np.flatnonzero(m)
    exp = sel[np.argsort(v[sel], kind="stable")]
    assert _ints(out, "v") == [int(x) for x in v[exp]]
    assert _ints(out, "w") == [int(x) for x in w[exp]]
    # bare OR as the whole WHERE, and an IN on a DATE column
    out = S.execute("SELECT COUNT(*) AS n FROM t WHERE v = 0 OR w = 49")
    assert _ints(out, "n") == [int(((v == 0) | (w == 49)).sum())]
    out = S.execute("SELECT w, COUNT(*) AS n, SUM(v) AS s FROM t "
                    "WHERE d IN (DATE '2000-01-01', DATE '2010-06-15', DATE '2030-12-31') "
                    "GROUP BY w ORDER BY w")
    import datetime
    days = [(datetime.date(y, mo, dd) - datetime.date(1970, 1, 1)).days
            for y, mo, dd in ((2000, 1, 1), (2010, 6, 15), (2030, 12, 31))]
    m = np.isin(d, days)
    g = pd.DataFrame({"w": w[m], "v": v[m]}).groupby("w").v.agg(["size", "sum"])
    assert _ints(out, "w") == list(g.index)
    assert _ints(out, "n") == [int(x) for x in g["size"]]
    assert _ints(out, "s") == [int(x) for x in g["sum"]]


def test_or_in_a_join_and_the_star_shape(ops):
    c, o, l = _star(ops, 7)
    out = S.execute("SELECT c.c_region, COUNT(*) AS n FROM lines l "
                    "JOIN orders o ON l.l_order = o.o_key JOIN cust c ON o.o_cust = c.c_key "
                    "WHERE c.c_region IN (1, 4, 6) AND (o.o_prio = 0 OR o.o_prio = 3) "
                    "GROUP BY c.c_region ORDER BY c.c_region")
    j = l.merge(o[o.o_prio.isin([0, 3])], left_on="l_order", right_on="o_key") \
        .merge(c[c.c_region.isin([1, 4, 6])], left_on="o_cust", right_on="c_key")
    g = j.groupby("c_region").size()
    assert _ints(out, "c_region") == list(g.index)
    assert _ints(out, "n") == [int(x) for x in g]


def test_multi_column_join_key(ops):
    """ON a.x = b.x AND a.y = b.y: the two columns are packed into one key
    on both sides; checked against pandas."""
    rng = np.random.default_rng(13)
    na, nb = 20_000, 400_000
    pairs = np.unique(rng.integers(0, 300, (na * 2, 2)), axis=0)[:na]
    rng.shuffle(pairs)
    ax, ay = pairs[:, 0].astype(np.int32), (pairs[:, 1] - 100).astype(np.int64)
    av = rng.integers(-10**6, 10**6, len(ax)).astype(np.int64)
    bx = rng.integers(0, 320, nb).astype(np.int32)
    by = (rng.integers(0, 300, nb) - 100).astype(np.int64)
    bw = rng.integers(0, 100, nb).astype(np.int32)
    _table(ops, "a", [("x", abi.INTEGER, ax), ("y", abi.BIGINT, ay), ("v", abi.BIGINT, av)])
    _table(ops, "b", [("x", abi.INTEGER, bx), ("y", abi.BIGINT, by), ("w", abi.INTEGER, bw)])
    out = S.execute("SELECT b.w, SUM(a.v) AS s, COUNT(*) AS n FROM a JOIN b "
                    "ON a.x = b.x AND b.y = a.y WHERE b.w < 60 GROUP BY b.w ORDER BY b.w")
    A = pd.DataFrame({"x": ax, "y": ay, "v": av})
    B = pd.DataFrame({"x": bx, "y": by, "w": bw})
    j = B[B.w < 60].merge(A, on=["x", "y"])
    g = j.groupby("w").v.agg(["sum", "size"])
    assert _ints(out, "w") == list(g.index)
    assert _ints(out, "s") == [int(x) for x in g["sum"]]
    assert _ints(out, "n") == [int(x) for x in g["size"]]



def test_aggregates_over_expressions_of_joined_columns(ops):
    """SUM/MIN/MAX/AVG of integer and DECIMAL expressions whose columns
    come from different joined tables (Q3's SUM(price * (100 - discount))
    once an IN list moves it off the star shape), against pandas."""
    c, o, l = _star(ops, 17)
    price = np.random.default_rng(5).integers(100, 100_000, len(l)).astype(np.int64)
    _table(ops, "items", [("l_order", abi.INTEGER, l.l_order.to_numpy()),
                          ("l_disc", abi.INTEGER, l.l_disc.to_numpy()),
                          ("l_qty", abi.BIGINT, l.l_qty.to_numpy()),
                          ("price", abi.DECIMAL, price, False, 2)])
    out = S.execute("SELECT c.c_region, SUM(l.price * (100 - l.l_disc)) AS rev, "
                    "MAX(l.l_qty - o.o_prio) AS hi, MIN(o.o_prio * 7 + l.l_disc) AS lo, "
                    "AVG(l.l_qty * 2) AS avg2, COUNT(*) AS n "
                    "FROM items l JOIN orders o ON l.l_order = o.o_key "
                    "JOIN cust c ON o.o_cust = c.c_key "
                    "WHERE c.c_region IN (0, 2, 5) GROUP BY c.c_region ORDER BY c.c_region")
    L = l.assign(price=price)
    j = L.merge(o, left_on="l_order", right_on="o_key") \
        .merge(c[c.c_region.isin([0, 2, 5])], left_on="o_cust", right_on="c_key")
    j["rev"] = j.price * (100 - j.l_disc)
    j["hi"] = j.l_qty - j.o_prio
    j["lo"] = j.o_prio * 7 + j.l_disc
    j["q2"] = j.l_qty * 2
    g = j.groupby("c_region").agg(rev=("rev", "sum"), hi=("hi", "max"), lo=("lo", "min"),
                                  q2=("q2", "mean"), n=("rev", "size"))
    assert _ints(out, "c_region") == list(g.index)
    t, sc, _ = out["rev"]
    assert (t, sc) == (abi.HUGEINT, 2) or (t, sc) == (abi.DECIMAL, 2)
    assert _ints(out, "rev") == [int(x) for x in g.rev]
    assert _ints(out, "hi") == [int(x) for x in g.hi]
    assert _ints(out, "lo") == [int(x) for x in g.lo]
    assert _ints(out, "n") == [int(x) for x in g.n]
    np.testing.assert_allclose(np.asarray(out["avg2"][2], np.float64), g.q2.to_numpy(), rtol=1e-12)


def test_expression_aggregate_overflow_is_an_error(ops):
    big = np.full(1000, 2**40, np.int64)
    _table(ops, "a", [("k", abi.INTEGER, np.arange(1000, dtype=np.int32), True),
                      ("v", abi.BIGINT, big)])
    _table(ops, "b", [("k", abi.INTEGER, np.arange(1000, dtype=np.int32))])
    with pytest.raises(abi.MettleError) as ei:
        S.execute("SELECT SUM(a.v * a.v) AS s FROM a JOIN b ON a.k = b.k")
    assert "Overflow" in str(ei.value)


def test_order_by_aggregate_then_group_keys(ops):
    """ORDER BY an aggregate, then group keys, LIMIT k (Q3's ORDER BY
    revenue DESC, o_orderdate): ties at the LIMIT boundary are ordered by the
    named keys, then by the remaining group keys ascending."""
    c, o, l = _star(ops, 19)
    j = l.merge(o, left_on="l_order", right_on="o_key").merge(c, left_on="o_cust", right_on="c_key")
    # many groups, heavy ties in the count: the top-K widens to every group
    out = S.execute("SELECT o.o_key, o.o_prio, COUNT(*) AS n, SUM(l.l_qty * l.l_disc) AS s "
                    "FROM lines l JOIN orders o ON l.l_order = o.o_key "
                    "GROUP BY o.o_key, o.o_prio ORDER BY n DESC, o.o_prio DESC, o.o_key LIMIT 20")
    g = j.assign(x=j.l_qty * j.l_disc).groupby(["o_key", "o_prio"]).agg(
        n=("x", "size"), s=("x", "sum")).reset_index()
    g = g.sort_values(["n", "o_prio", "o_key"], ascending=[False, False, True]).head(20)
    assert _ints(out, "o_key") == [int(x) for x in g.o_key]
    assert _ints(out, "o_prio") == [int(x) for x in g.o_prio]
    assert _ints(out, "n") == [int(x) for x in g.n]
    assert _ints(out, "s") == [int(x) for x in g.s]
    # few groups: the candidate set holds them all
    out = S.execute("SELECT c.c_region, o.o_prio, SUM(l.l_qty) AS q FROM lines l "
                    "JOIN orders o ON l.l_order = o.o_key JOIN cust c ON o.o_cust = c.c_key "
                    "GROUP BY o.o_prio, c.c_region ORDER BY q, c.c_region DESC LIMIT 7")
    g = j.groupby(["o_prio", "c_region"]).l_qty.sum().rename("q").reset_index()
    g = g.sort_values(["q", "c_region", "o_prio"], ascending=[True, False, True]).head(7)
    assert _ints(out, "c_region") == [int(x) for x in g.c_region]
    assert _ints(out, "o_prio") == [int(x) for x in g.o_prio]
    assert _ints(out, "q") == [int(x) for x in g.q]


def _profiled_kernels(fn):
    """Names of the library kernels fn() launched (the library's profiler)."""
    import ctypes as C
    L = abi.lib()
    L.mtl_profile_reset()
    L.mtl_profile_enable(1)
    try:
        out = fn()
        abi.check(L.mtl_stream_sync(None))  # resolves the recorded launches
    finally:
        L.mtl_profile_enable(0)
    names, i = [], 0
    name, ms, cnt = C.create_string_buffer(128), C.c_double(), C.c_uint64()
    while L.mtl_profile_get(i, name, 128, C.byref(ms), C.byref(cnt)) == 0:
        names.append(name.value.decode())
        i += 1
    L.mtl_profile_reset()
    return out, names


@pytest.mark.parametrize("nb", [5_000, 300_000])
def test_grouped_join_on_sparse_keys_uses_radix_partitions(ops, nb):
    """A grouped join whose build keys are too sparse for the dense table
    takes the LDS-staged radix partitions (pairs in partition order, which
    aggregation does not see); a projection keeps the probe-ordered hash
    join.  Both against pandas."""
    rng = np.random.default_rng(nb)
    npr = 1_000_000
    ak = (rng.choice(10**12, nb, replace=False) - 5 * 10**11).astype(np.int64)
    ag = rng.integers(0, 40, nb).astype(np.int32)
    bk = ak[rng.integers(0, nb, npr)]
    bk[rng.random(npr) < 0.2] = 10**12 + 7  # no match
    bv = rng.integers(-10**6, 10**6, npr).astype(np.int64)
    _table(ops, "a", [("k", abi.BIGINT, ak, True), ("g", abi.INTEGER, ag)])
    _table(ops, "b", [("k", abi.BIGINT, bk), ("v", abi.BIGINT, bv)])
    q = "SELECT a.g, COUNT(*) AS n, SUM(b.v) AS s FROM a JOIN b ON a.k = b.k GROUP BY a.g ORDER BY a.g"
    out, names = _profiled_kernels(lambda: S.execute(q))
    assert "k_rx_join" in names and "k_probe" not in names
    j = pd.DataFrame({"k": bk, "v": bv}).merge(pd.DataFrame({"k": ak, "g": ag}), on="k")
    g = j.groupby("g").v.agg(["size", "sum"])
    assert _ints(out, "g") == list(g.index)
    assert _ints(out, "n") == [int(x) for x in g["size"]]
    assert _ints(out, "s") == [int(x) for x in g["sum"]]
    out, names = _profiled_kernels(lambda: S.execute(
        "SELECT b.v, a.g FROM a JOIN b ON a.k = b.k WHERE b.v > 999000"))
    assert "k_rx_join" not in names
    m = j[j.v > 999000]
    assert sorted(zip(_ints(out, "v"), _ints(out, "g"))) == sorted(zip(m.v, m.g))
