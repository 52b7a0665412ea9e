"""SQL lowering without a GPU: the admitted subset parses, binds against the
catalog and picks the expected shape; everything else is refused with a
reason (the DuckDB CPU fallback); gpu_rewrite_query prints the hook form,
which plans to the same shape.  Catalog pointers are placeholders: planning
never dereferences them."""
import pytest

from mettle_amd import abi
from mettle_amd import sql as S

FAKE = 0x1000


@pytest.fixture(autouse=True)
def catalog():
    S.drop()
    S.register_table("numbers", 1_000_000, [("value", abi.INTEGER, FAKE, 0, 2**20 - 1, False)])
    S.register_table("t", 100, [("c_i", abi.INTEGER, FAKE, -2**31, 2**31 - 1, False),
                                ("c_b", abi.BIGINT, FAKE, -2**63, 2**63 - 1, False),
                                ("c_f", abi.FLOAT, FAKE, 0, 0, False),
                                ("c_d", abi.DATE, FAKE, 10957, 22280, False),
                                ("c_t", abi.TIMESTAMP, FAKE, 0, 2**62, False),
                                ("g", abi.BIGINT, FAKE, 0, 999, False),
                                ("price", abi.DECIMAL, FAKE, 0, 10**9, False, 2)])
    S.register_table("customers", 10, [("c_key", abi.INTEGER, FAKE, 1, 10, True),
                                       ("segment", abi.DICT8, FAKE, 0, 4, False)])
    S.set_dictionary("customers", "segment", ["AUTOMOBILE", "BUILDING", "FURNITURE",
                                              "HOUSEHOLD", "MACHINERY"])
    S.register_table("tags", 4, [("tag", abi.DICT8, FAKE, 0, 1, False)])
    S.set_dictionary("tags", "tag", ["zeta", "alpha"])  # not sorted
    S.register_table("orders", 100, [("o_key", abi.INTEGER, FAKE, 1, 100, True),
                                     ("o_cust", abi.INTEGER, FAKE, 1, 10, False),
                                     ("o_date", abi.DATE, FAKE, 8035, 10592, False),
                                     ("o_prio", abi.INTEGER, FAKE, 0, 4, False)])
    S.register_table("lines", 400, [("l_order", abi.INTEGER, FAKE, 1, 100, False),
                                    ("price", abi.DECIMAL, FAKE, 0, 10**8, False, 2),
                                    ("discount", abi.BIGINT, FAKE, 0, 10, False),
                                    ("l_ship", abi.DATE, FAKE, 8035, 10800, False)])
    yield
    S.drop()


Q1 = "SELECT value, value * 10 AS scaled FROM numbers WHERE value > 524288 ORDER BY value"
Q3 = ("SELECT l.l_order, SUM(l.price * (100 - l.discount)) AS revenue, o.o_date, o.o_prio "
      "FROM customers c, orders o, lines l WHERE c.segment = 'BUILDING' AND c.c_key = o.o_cust "
      "AND l.l_order = o.o_key AND o.o_date < DATE '1995-03-15' AND l.l_ship > DATE '1995-03-15' "
      "GROUP BY l.l_order, o.o_date, o.o_prio ORDER BY revenue DESC, o.o_date LIMIT 10")


@pytest.mark.parametrize("sql,shape", [
    (Q1, "scan"),
    ("SELECT * FROM t WHERE c_d BETWEEN DATE '2001-01-01' AND DATE '2002-01-01'", "scan"),
    ("SELECT c_b + 7 FROM t WHERE c_b < -9000000000000000000 LIMIT 5", "scan"),
    ("SELECT c_t FROM t WHERE c_t >= TIMESTAMP '2010-06-01 12:00:00' ORDER BY c_t DESC", "scan"),
    ("SELECT c_f * 2.5 FROM t WHERE 0.5 < c_f", "scan"),
    ("SELECT price FROM t WHERE price >= 12.25", "scan"),
    ("SELECT g, SUM(c_b), COUNT(*), MIN(c_b), MAX(c_b), AVG(c_f) FROM t GROUP BY g", "group"),
    ("SELECT g, SUM(c_b) AS s FROM t WHERE c_i > 0 GROUP BY g ORDER BY s DESC LIMIT 100",
     "group"),
    (Q3, "star3"),
    ("SELECT o.o_key, SUM(l.price) AS r FROM orders o JOIN lines l ON l.l_order = o.o_key "
     "GROUP BY o.o_key ORDER BY r DESC LIMIT 3", "star2"),
    # the generic joins: grouped by a non-key column, projections, ungrouped
    ("SELECT o.o_date, SUM(l.price) AS r FROM orders o, lines l WHERE l.l_order = o.o_key "
     "GROUP BY o.o_date ORDER BY r DESC LIMIT 3", "join"),
    ("SELECT g, SUM(c_i) FROM t GROUP BY g ORDER BY g DESC", "join"),
    ("SELECT SUM(value) AS total, COUNT(*) AS cnt, MIN(value), MAX(value) FROM numbers "
     "WHERE value >= 20", "join"),
    ("SELECT o.o_key, l.price FROM orders o JOIN lines l ON l.l_order = o.o_key "
     "WHERE o.o_prio = 2 ORDER BY l.price DESC LIMIT 10", "join"),
    ("SELECT c.segment, o.o_prio, COUNT(*), SUM(l.discount) FROM customers c "
     "JOIN orders o ON c.c_key = o.o_cust JOIN lines l ON l.l_order = o.o_key "
     "GROUP BY c.segment, o.o_prio ORDER BY c.segment, o.o_prio", "join"),
    ("SELECT COUNT(*) FROM orders o JOIN lines l ON l.l_order = o.o_key", "join"),
    # a sorted dictionary admits range comparisons
    ("SELECT segment FROM customers WHERE segment < 'B'", "scan"),
    # OR runs and IN lists
    ("SELECT value FROM numbers WHERE value > 1 OR value < 0", "scan"),
    ("SELECT value FROM numbers WHERE value IN (1, 5, 9) AND (value < 4 OR value > 8)", "scan"),
    ("SELECT g, COUNT(*) FROM t WHERE g IN (1, 3) OR c_i < 0 GROUP BY g", "group"),
    ("SELECT o.o_key FROM customers c JOIN orders o ON c.c_key = o.o_cust "
     "WHERE c.segment IN ('BUILDING', 'MACHINERY') AND (o.o_prio = 1 OR o.o_prio > 3)", "join"),
    (Q3.replace("c.segment = 'BUILDING'", "c.segment IN ('BUILDING', 'FURNITURE')"), "join"),
    # two equalities between one pair of tables: a packed two-column key
    ("SELECT o.o_key, l.price FROM orders o JOIN lines l ON l.l_order = o.o_key "
     "AND l.discount = o.o_prio", "join"),
])
def test_admitted_shapes(sql, shape):
    assert S.explain(sql) == shape


@pytest.mark.parametrize("sql,why", [
    ("SELECT value FROM numbers WHERE value > 1 AND value < 9 OR value = 0", "parentheses"),
    ("SELECT value FROM numbers WHERE value = 0 OR value > 1 AND value < 9", "parentheses"),
    ("SELECT value FROM numbers WHERE (value > 1 AND value < 9) OR value = 0", "OR over a conj"),
    ("SELECT value FROM numbers WHERE value BETWEEN 1 AND 9 OR value = 0", "OR over a conj"),
    ("SELECT value FROM numbers WHERE value NOT IN (1, 2)", "NOT"),
    ("SELECT o.o_key FROM customers c JOIN orders o ON c.c_key = o.o_cust "
     "WHERE c.c_key = 1 OR o.o_prio = 2", "OR across tables"),
    ("SELECT DISTINCT value FROM numbers", "DISTINCT"),
    ("SELECT value FROM numbers WHERE value / 2 > 1", "expected a comparison"),
    ("SELECT value / 2 FROM numbers", "division"),
    ("SELECT value FROM missing", "not registered"),
    ("SELECT c_b FROM t ORDER BY c_b, c_i", "more than one"),
    ("SELECT c_i FROM t WHERE c_i > 1.5", "non-integer literal"),
    ("SELECT c_d FROM t WHERE c_d < DATE '2001-02-30'", "invalid date"),
    ("SELECT g, SUM(c_b) FROM t GROUP BY g HAVING SUM(c_b) > 0", "HAVING"),
    ("SELECT tag FROM tags WHERE tag < 'b'", "not sorted"),
    ("SELECT price FROM t WHERE price > 1.255", "more scale"),
    ("SELECT o.o_key, l.price FROM orders o, lines l", "tree"),
    ("SELECT o.o_key, l.price FROM orders o JOIN lines l ON l.l_order = o.o_key "
     "ORDER BY o.o_key, l.price", "more than one"),
    ("SELECT SUM(c_b), c_i FROM t", "neither grouped"),
    ("SELECT o.o_key, SUM(l.price) AS r FROM orders o, lines l WHERE l.l_order = o.o_key "
     "GROUP BY o.o_date ORDER BY r DESC LIMIT 3", "neither grouped"),
    ("SELECT c_b, SUM(c_i) FROM t GROUP BY c_b ORDER BY c_b", "hashed"),
    ("SELECT o.o_prio, SUM(l.price) FROM orders o JOIN lines l ON l.l_order = o.o_key "
     "GROUP BY o.o_prio ORDER BY SUM(l.discount)", "ORDER BY of an aggregate"),
    ("SELECT c_b FROM t WHERE c_b = 1; SELECT 1", "unexpected"),
    ("SELECT g, SUM(c_b) AS s FROM t GROUP BY g ORDER BY s DESC LIMIT 4294967296",
     "exceeds 2^32-1"),
    ("SELECT g, SUM(c_b) AS s FROM t GROUP BY g ORDER BY s DESC LIMIT 4294967297",
     "exceeds 2^32-1"),
])
def test_refused_with_reason(sql, why):
    with pytest.raises(abi.Unsupported) as ei:
        S.explain(sql)
    assert why in str(ei.value)


def test_rewrite_uses_hooks_and_round_trips():
    r = S.rewrite(Q1)
    assert r == ("SELECT value, gpu_mul_scalar(value, 10) AS scaled FROM numbers "
                 "WHERE gpu_gt_scalar(value, 524288) ORDER BY value")
    assert S.explain(r) == "scan"
    r3 = S.rewrite(Q3)
    assert "gpu_eq_scalar(c.segment, 'BUILDING')" in r3
    assert "gpu_lt_scalar(o.o_date, DATE '1995-03-15')" in r3
    assert "c.c_key = o.o_cust" in r3
    assert S.explain(r3) == "star3"
    # literal on the left flips the comparison
    assert "gpu_gt_scalar(c_f, 0.5)" in S.rewrite("SELECT c_f FROM t WHERE 0.5 < c_f")


def test_rewrite_refuses_outside_the_subset():
    with pytest.raises(abi.Unsupported):
        S.rewrite("SELECT a FROM t WHERE a NOT IN (1, 2)")
