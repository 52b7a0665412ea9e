"""Pins the CPU oracle before its results are trusted as expected values.

Each function of oracle/mettle_oracle.c is checked against something it does
not restate: the reference's own fixtures (DuckDB-produced rows in
test/sql/*.test, hand-computed expectations in tests/extension.rs, the date
unit tests in src/plan/mod.rs), numpy closed forms, and brute force on tiny
inputs.  A planted mistake (wrong op, dropped predicate, swapped operand,
wrong tie order) fails at least one of these.
"""
import math

import numpy as np
import pytest

from mettle_amd import abi, datagen
from oracle import oracle as orc
from tests import golden

NP_OPS = {abi.CMP_EQ: np.equal, abi.CMP_LT: np.less, abi.CMP_LE: np.less_equal,
          abi.CMP_GT: np.greater, abi.CMP_GE: np.greater_equal}


# ---------------------------------------------------------------------------
# predicates + compaction
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("t,dtype", [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64),
                                     (abi.DATE, np.int32), (abi.TIMESTAMP, np.int64),
                                     (abi.DICT8, np.uint8), (abi.FLOAT, np.float32),
                                     (abi.DOUBLE, np.float64)])
@pytest.mark.parametrize("op", list(NP_OPS))
def test_filter_matches_numpy_closed_form(t, dtype, op):
    rng = np.random.default_rng(7 + t * 10 + op)
    n = 5000
    if np.issubdtype(dtype, np.floating):
        col = rng.standard_normal(n).astype(dtype)
        lit = float(col[17])
    else:
        info = np.iinfo(dtype)
        col = rng.integers(max(info.min, -1000), min(info.max, 1000), n).astype(dtype)
        lit = int(col[17])
    got = orc.filter([(col, t, op, lit)], n)
    ref = np.flatnonzero(NP_OPS[op](col, dtype(lit))).astype(np.uint32)
    np.testing.assert_array_equal(got, ref)


def test_filter_float_nan_total_order():
    # DuckDB: NaN equals NaN and is greater than every other value
    col = np.array([np.nan, 1.0, -np.inf, np.inf, np.nan], dtype=np.float64)
    assert list(orc.filter([(col, abi.DOUBLE, abi.CMP_EQ, float("nan"))], 5)) == [0, 4]
    assert list(orc.filter([(col, abi.DOUBLE, abi.CMP_GT, float("inf"))], 5)) == [0, 4]
    assert list(orc.filter([(col, abi.DOUBLE, abi.CMP_LT, float("nan"))], 5)) == [1, 2, 3]


def test_filter_conjunction_and_selection_refinement():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 100, 2000).astype(np.int32)
    b = rng.integers(0, 100, 2000).astype(np.int64)
    got = orc.filter([(a, abi.INTEGER, abi.CMP_GT, 30), (b, abi.BIGINT, abi.CMP_LE, 60)], 2000)
    ref = np.flatnonzero((a > 30) & (b <= 60)).astype(np.uint32)
    np.testing.assert_array_equal(got, ref)
    sel = orc.filter([(a, abi.INTEGER, abi.CMP_GT, 30)], 2000)
    got2 = orc.filter([(b, abi.BIGINT, abi.CMP_LE, 60)], 2000, in_sel=sel)
    np.testing.assert_array_equal(got2, ref)


def test_filter_reference_fixtures():
    # tests/extension.rs:61-92: [1..5] WHERE value > 2 -> [3,4,5]
    v = np.array([1, 2, 3, 4, 5], dtype=np.int32)
    assert list(v[orc.filter([(v, abi.INTEGER, abi.CMP_GT, 2)], 5)]) == [3, 4, 5]
    # sirius-compat.test:48-52 BIGINT WHERE value >= 20 -> 20, 30
    b = np.array([10, 20, 30], dtype=np.int64)
    assert list(b[orc.filter([(b, abi.BIGINT, abi.CMP_GE, 20)], 3)]) == [20, 30]
    # sirius-compat.test:58-62 FLOAT WHERE value >= 2.0 -> 2.0, 3.5
    f = np.array([1.0, 2.0, 3.5], dtype=np.float32)
    assert list(f[orc.filter([(f, abi.FLOAT, abi.CMP_GE, 2.0)], 3)]) == [2.0, 3.5]
    # sirius-compat.test:71-75 DATE >= '1995-03-25'
    d = np.array([orc.date_days(1995, 3, 1), orc.date_days(1995, 3, 25),
                  orc.date_days(1995, 4, 1)], dtype=np.int32)
    got = orc.filter([(d, abi.DATE, abi.CMP_GE, orc.date_days(1995, 3, 25))], 3)
    assert list(got) == [1, 2]
    # sirius-compat.test:84-88 TIMESTAMP > '1995-03-01 12:00:00'
    ts = np.array([golden.ts_us(1995, 3, 1, 8, 0, 0), golden.ts_us(1995, 3, 1, 12, 34, 56),
                   golden.ts_us(1995, 4, 1, 1, 2, 3)], dtype=np.int64)
    got = orc.filter([(ts, abi.TIMESTAMP, abi.CMP_GT, golden.ts_us(1995, 3, 1, 12, 0, 0))], 3)
    assert list(got) == [1, 2]


@pytest.mark.parametrize("op", list(NP_OPS))
@pytest.mark.parametrize("lit", [2**40, -2**40, 2**31, -2**31 - 1])
def test_filter_int_column_wide_literal(op, lit):
    """An INTEGER column against a literal outside int32: compared in int64
    (truncating the literal to 32 bits would flip these results)."""
    v = np.array([-2**31, -5, 0, 5, 2**31 - 1], dtype=np.int32)
    got = orc.filter([(v, abi.INTEGER, op, lit)], len(v))
    ref = np.flatnonzero(NP_OPS[op](v.astype(np.int64), np.int64(lit))).astype(np.uint32)
    np.testing.assert_array_equal(got, ref)


def test_filter_planted_op_mistake_is_caught():
    # swapping LT for LE changes the result on the boundary literal
    v = np.array([1, 2, 3], dtype=np.int32)
    assert list(orc.filter([(v, abi.INTEGER, abi.CMP_LT, 2)], 3)) == [0]
    assert list(orc.filter([(v, abi.INTEGER, abi.CMP_LE, 2)], 3)) == [0, 1]


# ---------------------------------------------------------------------------
# projections
# ---------------------------------------------------------------------------
def test_project_fixture_values():
    v = np.array([1, 2, 3], dtype=np.int32)
    # sirius-compat.test:18-29: value*10 -> 10,20,30 ; value+5 -> 6,7,8
    assert list(orc.project(v, abi.INTEGER, None, abi.PROJ_MUL, 10)[0]) == [10, 20, 30]
    assert list(orc.project(v, abi.INTEGER, None, abi.PROJ_ADD, 5)[0]) == [6, 7, 8]
    f = np.array([2.0, 3.5], dtype=np.float32)
    # sirius-compat.test:58-62: value*1.5 -> 3.0, 5.25
    assert list(orc.project(f, abi.FLOAT, None, abi.PROJ_MUL, 1.5)[0]) == [3.0, 5.25]
    assert list(orc.project(v, abi.INTEGER, None, abi.PROJ_SUBFROM, 10)[0]) == [9, 8, 7]
    assert list(orc.project(v, abi.INTEGER, None, abi.PROJ_SUB, 10)[0]) == [-9, -8, -7]


@pytest.mark.parametrize("op", [abi.PROJ_ADD, abi.PROJ_MUL, abi.PROJ_SUB, abi.PROJ_SUBFROM])
def test_project_int32_matches_wide_arithmetic(op):
    rng = np.random.default_rng(op)
    v = rng.integers(-(1 << 31), 1 << 31, 4000, dtype=np.int64).astype(np.int32)
    s = 12345
    got, ovf = orc.project(v, abi.INTEGER, None, op, s)
    w = v.astype(np.int64)
    ref = {abi.PROJ_ADD: w + s, abi.PROJ_MUL: w * s, abi.PROJ_SUB: w - s,
           abi.PROJ_SUBFROM: s - w}[op]
    assert ovf == bool(((ref < -(1 << 31)) | (ref >= (1 << 31))).any())
    np.testing.assert_array_equal(got, ref.astype(np.int32))  # wrapped bits


def test_project_bigint_overflow_flag():
    v = np.array([np.iinfo(np.int64).max - 1, 5], dtype=np.int64)
    _, ovf = orc.project(v, abi.BIGINT, None, abi.PROJ_ADD, 1)
    assert not ovf
    _, ovf = orc.project(v, abi.BIGINT, None, abi.PROJ_ADD, 2)
    assert ovf


def test_project_float_is_ieee_f32():
    rng = np.random.default_rng(11)
    f = rng.standard_normal(3000).astype(np.float32)
    got, _ = orc.project(f, abi.FLOAT, None, abi.PROJ_MUL, 1.1)
    np.testing.assert_array_equal(got.view(np.uint32), (f * np.float32(1.1)).view(np.uint32))


# ---------------------------------------------------------------------------
# joins
# ---------------------------------------------------------------------------
def brute_join(bk, pk):
    pairs = [(p, b) for p in range(len(pk)) for b in range(len(bk)) if pk[p] == bk[b]]
    return [p for p, _ in pairs], [b for _, b in pairs]


def test_join_matches_nested_loop():
    rng = np.random.default_rng(5)
    bk = rng.permutation(np.arange(-50, 150)).astype(np.int64)[:120]
    pk = rng.integers(-80, 200, 400).astype(np.int64)
    p, b = orc.join(bk, abi.BIGINT, pk, abi.BIGINT)
    rp, rb = brute_join(bk, pk)
    assert list(p) == rp and list(b) == rb


def test_join_reference_fixture_two_table():
    # tests/extension.rs:279-317: (10,100),(11,200) from a PK/FK join
    orders_cust = np.array([1, 2], dtype=np.int32)
    cust_key = np.array([2, 1], dtype=np.int32)
    p, b = orc.join(cust_key, abi.INTEGER, orders_cust, abi.INTEGER)
    assert list(p) == [0, 1] and list(b) == [1, 0]


def test_join_duplicate_build_key_is_unsupported():
    bk = np.array([1, 2, 2], dtype=np.int32)
    with pytest.raises(abi.Unsupported):
        orc.join(bk, abi.INTEGER, bk, abi.INTEGER)


# ---------------------------------------------------------------------------
# star join -> grouped SUM -> top-K
# ---------------------------------------------------------------------------
def star_desc_host(tables, **kw):
    cust, orders, lines = tables
    by = {"customers": cust, "orders": orders, "lines": lines}
    d = datagen.star_desc(lambda t, c: by[t][c].ctypes.data, len(cust["c_key"]),
                          len(orders["o_key"]), len(lines["l_order"]), **kw)
    d.d_key_min, d.d_key_max = int(cust["c_key"].min()), int(cust["c_key"].max())
    d.p_key_min, d.p_key_max = int(orders["o_key"].min()), int(orders["o_key"].max())
    return d


def brute_star(tables, segment, cutoff, k):
    cust, orders, lines = tables
    ok_c = {int(c) for c, s in zip(cust["c_key"], cust["segment"]) if s == segment}
    prow = {}
    for i in range(len(orders["o_key"])):
        if orders["o_date"][i] < cutoff and int(orders["o_cust"][i]) in ok_c:
            prow[int(orders["o_key"][i])] = i
    acc = {}
    for i in range(len(lines["l_order"])):
        key = int(lines["l_order"][i])
        if key in prow and lines["l_ship"][i] > cutoff:
            v = int(lines["price"][i]) * (100 - int(lines["discount"][i]))
            s, c = acc.get(key, (0, 0))
            acc[key] = (s + v, c + 1)
    groups = [(-s, int(orders["o_date"][prow[key]]), key, c) for key, (s, c) in acc.items()]
    groups.sort()
    return [(prow[key], -ns, c) for ns, _, key, c in groups[:k]]


def test_star_matches_brute_force_small_random():
    tables = datagen.star_tables(scale=0.0004)  # 600 customers, 6000 orders
    cutoff = datagen.STAR_CUTOFF
    for seg in range(5):
        d = star_desc_host(tables, segment=seg, cutoff=cutoff, k=25)
        st, rows = orc.star_topk(d)
        assert st == abi.MTL_OK
        got = [(r.p_row, r.sum_i64, r.count) for r in rows]
        assert got == brute_star(tables, seg, cutoff, 25)


def test_star_ties_break_on_date_then_key():
    # equal revenues: order by o_date asc, then key asc
    cust = {"c_key": np.array([1], np.int32), "segment": np.array([1], np.uint8)}
    orders = {"o_key": np.array([10, 11, 12, 13], np.int32),
              "o_cust": np.array([1, 1, 1, 1], np.int32),
              "o_date": np.array([5, 3, 3, 9], np.int32),
              "o_prio": np.zeros(4, np.int32)}
    lines = {"l_order": np.array([10, 11, 12, 13], np.int32),
             "price": np.array([100, 100, 100, 200], np.int64),
             "discount": np.zeros(4, np.int64),
             "l_ship": np.array([50, 50, 50, 50], np.int32)}
    d = star_desc_host((cust, orders, lines), segment=1, cutoff=20, k=4)
    st, rows = orc.star_topk(d)
    assert st == abi.MTL_OK
    assert [r.p_row for r in rows] == [3, 1, 2, 0]


def test_star_reference_q3_fixture_integer_cents():
    st, rows, _ = golden.q3_fixture_run(orc.star_topk, decimal=True)
    assert st == abi.MTL_OK
    # tpch-sirius.test:58-61 -> 101 140.0 1995-03-01 1 ; 104 27.0 1995-03-10 2
    assert [(r.p_row, r.sum_i64) for r in rows] == [(0, 1_400_000), (3, 270_000)]


def test_star_reference_q3_fixture_double():
    st, rows, _ = golden.q3_fixture_run(orc.star_topk, decimal=False)
    assert st == abi.MTL_OK
    assert [(r.p_row, r.sum_f64) for r in rows] == [(0, 140.0), (3, 27.0)]


def test_star_boundary_rows_are_excluded():
    # tpch-sirius.test:55-61 adds order 105 (o_date = cutoff) and 106
    # (l_shipdate = cutoff): strict < and > keep them out.
    st, rows, _ = golden.q3_fixture_run(orc.star_topk, decimal=True, with_boundary_rows=True)
    assert st == abi.MTL_OK
    assert [(r.p_row, r.sum_i64) for r in rows] == [(0, 1_400_000), (3, 270_000)]


def test_star_duplicate_parent_key_is_unsupported():
    cust = {"c_key": np.array([1], np.int32), "segment": np.array([1], np.uint8)}
    orders = {"o_key": np.array([10, 10], np.int32), "o_cust": np.array([1, 1], np.int32),
              "o_date": np.array([1, 1], np.int32), "o_prio": np.zeros(2, np.int32)}
    lines = {"l_order": np.array([10], np.int32), "price": np.array([1], np.int64),
             "discount": np.zeros(1, np.int64), "l_ship": np.array([50], np.int32)}
    d = star_desc_host((cust, orders, lines), segment=1, cutoff=20, k=4)
    st, _ = orc.star_topk(d)
    assert st == abi.MTL_UNSUPPORTED


# ---------------------------------------------------------------------------
# date literals
# ---------------------------------------------------------------------------
def test_date_days_unit_tests_from_planner():
    # src/plan/mod.rs:182-194
    assert orc.date_days(2024, 1, 1) == 19723
    assert orc.date_days(1970, 1, 1) == 0
    assert orc.date_days(1969, 12, 31) == -1
    assert orc.date_days(2000, 3, 1) - orc.date_days(2000, 2, 28) == 2  # leap year


def test_date_days_matches_python_calendar():
    import datetime
    epoch = datetime.date(1970, 1, 1).toordinal()
    rng = np.random.default_rng(1)
    for days in rng.integers(1 - epoch, datetime.date(9999, 12, 31).toordinal() - epoch, 2000):
        dt = datetime.date.fromordinal(int(days) + epoch)
        assert orc.date_days(dt.year, dt.month, dt.day) == days


def test_fsum_reference_for_double_sums():
    # the oracle's compensated double sum agrees with math.fsum
    cust = {"c_key": np.array([1], np.int32), "segment": np.array([1], np.uint8)}
    rng = np.random.default_rng(9)
    n = 20000
    vals = rng.standard_normal(n) * 1e6
    orders = {"o_key": np.array([1], np.int32), "o_cust": np.array([1], np.int32),
              "o_date": np.array([1], np.int32), "o_prio": np.zeros(1, np.int32)}
    lines = {"l_order": np.ones(n, np.int32), "price": vals, "discount": np.zeros(n),
             "l_ship": np.full(n, 50, np.int32)}
    d = star_desc_host((cust, orders, lines), segment=1, cutoff=20, k=1)
    d.agg = abi.make_expr([(abi.EXPR_COL, 0), (abi.EXPR_CONST, 1.0), (abi.EXPR_COL, 1),
                           (abi.EXPR_SUB, 0), (abi.EXPR_MUL, 0)],
                          [(lines["price"].ctypes.data, abi.DOUBLE),
                           (lines["discount"].ctypes.data, abi.DOUBLE)], is_double=True)
    st, rows = orc.star_topk(d)
    assert st == abi.MTL_OK
    assert math.isclose(rows[0].sum_f64, math.fsum(vals), rel_tol=0, abs_tol=1e-6)
