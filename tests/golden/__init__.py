"""Golden fixtures transcribed from the reference's tests.

Rows below are data copied from the DuckDB-produced expectations of
test/sql/tpch-sirius.test:7-61 and test/sql/sirius-compat.test:84-88,
111-161 (those fixtures run through DuckDB's CPU fallback, so their expected
rows are DuckDB output).  Dates are stored as days since 1970-01-01 and
timestamps as microseconds, computed with Python's calendar (not with the
oracle under test).
"""
import datetime

import numpy as np

from mettle_amd import abi

_EPOCH = datetime.date(1970, 1, 1).toordinal()


def days(y, m, d):
    return datetime.date(y, m, d).toordinal() - _EPOCH


def ts_us(y, mo, d, h, mi, s):
    return (days(y, mo, d) * 86400 + h * 3600 + mi * 60 + s) * 1_000_000


# c_mktsegment dictionary: HOUSEHOLD -> 0, AUTOMOBILE -> 1
Q3_CUSTOMER = [(1, 0), (2, 1), (3, 0)]
Q3_ORDERS = [(101, 1, days(1995, 3, 1), 1), (102, 2, days(1995, 3, 1), 2),
             (103, 3, days(1995, 4, 1), 3), (104, 3, days(1995, 3, 10), 2)]
Q3_ORDERS_BOUNDARY = [(105, 1, days(1995, 3, 25), 4), (106, 1, days(1995, 3, 24), 5)]
# (l_orderkey, l_extendedprice, l_discount, l_shipdate)
Q3_LINEITEM = [(101, 100.0, 0.10, days(1995, 4, 1)), (101, 50.0, 0.00, days(1995, 4, 3)),
               (102, 80.0, 0.20, days(1995, 4, 1)), (103, 70.0, 0.00, days(1995, 4, 10)),
               (104, 40.0, 0.25, days(1995, 2, 1)), (104, 30.0, 0.10, days(1995, 4, 10))]
Q3_LINEITEM_BOUNDARY = [(105, 500.0, 0.00, days(1995, 3, 26)), (106, 60.0, 0.00, days(1995, 3, 25))]
Q3_CUTOFF = days(1995, 3, 25)
# expected (l_orderkey, revenue, o_orderdate, o_shippriority), in order
Q3_EXPECTED = [(101, 140.0, days(1995, 3, 1), 1), (104, 27.0, days(1995, 3, 10), 2)]


def q3_tables(decimal: bool, with_boundary_rows: bool = False):
    """numpy columns named like the star workload.  decimal=True stores
    price in cents and discount in hundredths (DECIMAL as scaled int64)."""
    orders = Q3_ORDERS + (Q3_ORDERS_BOUNDARY if with_boundary_rows else [])
    items = Q3_LINEITEM + (Q3_LINEITEM_BOUNDARY if with_boundary_rows else [])
    cust = {"c_key": np.array([c for c, _ in Q3_CUSTOMER], np.int32),
            "segment": np.array([s for _, s in Q3_CUSTOMER], np.uint8)}
    o = {"o_key": np.array([r[0] for r in orders], np.int32),
         "o_cust": np.array([r[1] for r in orders], np.int32),
         "o_date": np.array([r[2] for r in orders], np.int32),
         "o_prio": np.array([r[3] for r in orders], np.int32)}
    if decimal:
        price = np.array([round(r[1] * 100) for r in items], np.int64)
        disc = np.array([round(r[2] * 100) for r in items], np.int64)
    else:
        price = np.array([r[1] for r in items], np.float64)
        disc = np.array([r[2] for r in items], np.float64)
    li = {"l_order": np.array([r[0] for r in items], np.int32), "price": price,
          "discount": disc, "l_ship": np.array([r[3] for r in items], np.int32)}
    return cust, o, li


def q3_desc(ptr, tables, decimal: bool, k: int = 10) -> abi.StarDesc:
    """Star descriptor of the Q3 fixture query: segment = HOUSEHOLD,
    o_orderdate < cutoff < l_shipdate, SUM(price * (1 - discount)),
    ORDER BY revenue DESC, o_orderdate.  ptr(table, col) -> address."""
    from mettle_amd import datagen
    cust, o, li = tables
    d = datagen.star_desc(ptr, len(cust["c_key"]), len(o["o_key"]), len(li["l_order"]), k=k,
                          segment=0, cutoff=Q3_CUTOFF)
    d.d_key_min, d.d_key_max = int(cust["c_key"].min()), int(cust["c_key"].max())
    d.p_key_min, d.p_key_max = int(o["o_key"].min()), int(o["o_key"].max())
    if not decimal:
        d.agg = abi.make_expr([(abi.EXPR_COL, 0), (abi.EXPR_CONST, 1.0), (abi.EXPR_COL, 1),
                               (abi.EXPR_SUB, 0), (abi.EXPR_MUL, 0)],
                              [(ptr("lines", "price"), abi.DOUBLE),
                               (ptr("lines", "discount"), abi.DOUBLE)], is_double=True)
    return d


def q3_fixture_run(star_fn, decimal: bool, with_boundary_rows: bool = False):
    """Runs star_fn (host pointers) on the fixture; returns (status, rows, tables)."""
    tables = q3_tables(decimal, with_boundary_rows)
    by = dict(zip(["customers", "orders", "lines"], tables))
    d = q3_desc(lambda t, c: by[t][c].ctypes.data, tables, decimal)
    st, rows = star_fn(d)
    return st, rows, tables
