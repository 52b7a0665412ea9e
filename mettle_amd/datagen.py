"""Seeded synthetic tables for the benchmark workloads.

numpy PCG64 with seed 20261203 + workload number.  Column widths follow the
declared DuckDB types: INTEGER/DATE int32, BIGINT/TIMESTAMP/DECIMAL int64,
FLOAT float32, dictionary codes uint8.
"""
from __future__ import annotations

import numpy as np

from . import abi

SEED_BASE = 20261203
DAYS_1992_01_01 = 8035      # days since 1970-01-01
DAYS_2000_01_01 = 10957
DAYS_2030_12_31 = 22279
SPAN_7Y = 2557              # o_date span in days


def rng_for(workload: int) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(SEED_BASE + workload))


def numbers_table(n: int = 1_000_000):
    """numbers(value INTEGER) uniform in [0, 2^20) (1M-row plain path)."""
    rng = rng_for(1)
    return {"value": rng.integers(0, 1 << 20, n, dtype=np.int32)}


def scan_table(n: int = 100_000_000, seed_offset: int = 0):
    """100M-row scan-filter-project table (c_i, c_b, c_f, c_d, c_t)."""
    rng = rng_for(2 + seed_offset)
    t = {
        "c_i": rng.integers(-(1 << 31), 1 << 31, n, dtype=np.int64).astype(np.int32),
        "c_b": rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n,
                            dtype=np.int64, endpoint=True),
        "c_f": rng.standard_normal(n, dtype=np.float32),
        "c_d": rng.integers(DAYS_2000_01_01, DAYS_2030_12_31 + 1, n, dtype=np.int32),
        "c_t": rng.integers(946684800 * 10**6, 1924991999 * 10**6 + 1, n, dtype=np.int64),
    }
    return t


def star_tables(scale: float = 1.0, workload: int = 3):
    """customers / orders / lines of the join -> group-agg -> top-10 workload.

    customers 1.5M x scale (c_key 1..N, segment u8 in 0..4), orders 15M x scale
    (o_key 1..N, o_cust FK, o_date over 7 years from 1992-01-01, o_prio 0..4),
    lines 1..7 per order clustered by order (l_order FK, price int64 cents in
    [90000, 10500000], discount int64 in 0..10, l_ship = o_date + U[1,121]).
    """
    rng = rng_for(workload)
    n_c = max(1, int(round(1_500_000 * scale)))
    n_o = max(1, int(round(15_000_000 * scale)))
    cust = {
        "c_key": np.arange(1, n_c + 1, dtype=np.int32),
        "segment": rng.integers(0, 5, n_c, dtype=np.uint8),
    }
    o_date = (DAYS_1992_01_01 + rng.integers(0, SPAN_7Y, n_o)).astype(np.int32)
    orders = {
        "o_key": np.arange(1, n_o + 1, dtype=np.int32),
        "o_cust": rng.integers(1, n_c + 1, n_o, dtype=np.int32),
        "o_date": o_date,
        "o_prio": rng.integers(0, 5, n_o, dtype=np.int32),
    }
    per = rng.integers(1, 8, n_o).astype(np.int64)
    n_l = int(per.sum())
    lines = {
        "l_order": np.repeat(orders["o_key"], per),
        "price": rng.integers(90000, 10500001, n_l, dtype=np.int64),
        "discount": rng.integers(0, 11, n_l, dtype=np.int64),
        "l_ship": (np.repeat(o_date, per) + rng.integers(1, 122, n_l)).astype(np.int32),
    }
    return cust, orders, lines


STAR_SEGMENT = 1
STAR_CUTOFF = DAYS_1992_01_01 + SPAN_7Y // 2

STAR_SQL = (
    "SELECT l.l_order, SUM(l.price * (100 - l.discount)) AS revenue, o.o_date, o.o_prio "
    "FROM customers c, orders o, lines l "
    "WHERE c.segment = 1 AND c.c_key = o.o_cust AND l.l_order = o.o_key "
    f"AND o.o_date < {STAR_CUTOFF} AND l.l_ship > {STAR_CUTOFF} "
    "GROUP BY l.l_order, o.o_date, o.o_prio ORDER BY revenue DESC, o.o_date LIMIT 10")


def star_desc(ptr, n_c: int, n_o: int, n_l: int, k: int = 10,
              segment: int = STAR_SEGMENT, cutoff: int = STAR_CUTOFF) -> abi.StarDesc:
    """Descriptor of the workload query.  ptr(table, column) -> address
    (host for the oracle, device for the HIP library)."""
    d = abi.StarDesc()
    d.d_rows = n_c
    d.d_key = ptr("customers", "c_key")
    d.d_key_type = abi.INTEGER
    d.d_key_min, d.d_key_max = 1, n_c
    d.d_n_preds = 1
    d.d_preds[0] = abi.make_pred(ptr("customers", "segment"), abi.DICT8, abi.CMP_EQ, segment)
    d.p_rows = n_o
    d.p_key = ptr("orders", "o_key")
    d.p_key_type = abi.INTEGER
    d.p_key_min, d.p_key_max = 1, n_o
    d.p_fk = ptr("orders", "o_cust")
    d.p_fk_type = abi.INTEGER
    d.p_n_preds = 1
    d.p_preds[0] = abi.make_pred(ptr("orders", "o_date"), abi.DATE, abi.CMP_LT, cutoff)
    d.f_rows = n_l
    d.f_fk = ptr("lines", "l_order")
    d.f_fk_type = abi.INTEGER
    d.f_n_preds = 1
    d.f_preds[0] = abi.make_pred(ptr("lines", "l_ship"), abi.DATE, abi.CMP_GT, cutoff)
    # price * (100 - discount), exact int64
    d.agg = abi.make_expr(
        [(abi.EXPR_COL, 0), (abi.EXPR_CONST, 100), (abi.EXPR_COL, 1), (abi.EXPR_SUB, 0),
         (abi.EXPR_MUL, 0)],
        [(ptr("lines", "price"), abi.DECIMAL), (ptr("lines", "discount"), abi.BIGINT)])
    d.k = k
    d.n_ties = 1
    d.tie_col[0] = ptr("orders", "o_date")
    d.tie_type[0] = abi.DATE
    d.tie_desc[0] = 0
    return d


# Algorithmic bytes per row (declared widths of every referenced column).
STAR_BYTES_PER_ROW = {"customers": 4 + 1, "orders": 4 + 4 + 4 + 4, "lines": 4 + 8 + 8 + 4}


def star_algorithmic_bytes(n_c: int, n_o: int, n_l: int) -> int:
    b = STAR_BYTES_PER_ROW
    return n_c * b["customers"] + n_o * b["orders"] + n_l * b["lines"]


def star_counts(cust, orders, lines, segment: int = STAR_SEGMENT,
                cutoff: int = STAR_CUTOFF) -> dict:
    """Row counts at each stage of the star query (numpy, for the per-kernel
    byte figures): qualifying customers, qualifying orders, lines whose
    order qualifies, and lines that also pass l_ship > cutoff."""
    c_ok = np.zeros(int(cust["c_key"].max()) + 1, dtype=bool)
    c_ok[cust["c_key"][cust["segment"] == segment]] = True
    o_pass = (orders["o_date"] < cutoff) & c_ok[orders["o_cust"]]
    o_ok = np.zeros(int(orders["o_key"].max()) + 1, dtype=bool)
    o_ok[orders["o_key"][o_pass]] = True
    l_parent = o_ok[lines["l_order"]]
    hits = l_parent & (lines["l_ship"] > cutoff)
    return {"dim_pass": int(c_ok.sum()), "parent_pass": int(o_pass.sum()),
            "fact_parent_rows": int(l_parent.sum()), "hits": int(hits.sum())}


def star_kernel_bytes(n_c: int, n_o: int, n_l: int, counts: dict) -> dict:
    """Algorithmic HBM bytes of each star-pipeline kernel per launch: every
    column a kernel must read, at its declared width, over the rows it must
    read it for, plus its outputs.  Late-materialised columns count only
    over the rows that reach them."""
    dp, pp = counts["dim_pass"], counts["parent_pass"]
    fp, h = counts["fact_parent_rows"], counts["hits"]
    return {
        # c_key 4 + segment 1; dimension bitmap and pass bits written
        "k_mark_dim": n_c * 5 + n_c // 8 * 2,
        # o_date 4 + o_cust 4 + o_key 4; parent bitmap and pass bits written
        "k_filter_parent": n_o * 12 + n_o // 8 * 2,
        # three bitmaps read, ranks written (one u32 per 64 parent keys)
        "k_rank": n_o // 8 * 2 + n_c // 8 + n_o // 64 * 4,
        # pass bits read; row_of (4 B) written and accumulators (12 B) cleared
        "k_map": n_o // 8 + pp * 16,
        # l_order 4 + l_ship 4 over all lines (both streamed in one pass),
        # hit row + slot (8 B) written
        "k_fact_scan": n_l * 8 + h * 8,
        # hit row + slot read, price 8 + discount 8, accumulators updated
        "k_fact_acc": h * (8 + 16) + min(h, pp) * 12,
        "k_star_topk": pp * 12,
        "k_star_topk_merge": 0,
    }


def group_table(n: int = 200_000_000, n_keys: int = 10_000_000, s: float = 1.1,
                dense_keys: int | None = None, workload: int = 4):
    """Skewed grouped-aggregate table: g BIGINT Zipf(s) over 1..n_keys by
    inverse-CDF sampling of the truncated distribution (or uniform over
    0..dense_keys-1), v_i BIGINT uniform in [-1e9, 1e9], v_f FLOAT N(0,1)."""
    rng = rng_for(workload)
    if dense_keys:
        g = rng.integers(0, dense_keys, n, dtype=np.int64)
    else:
        w = 1.0 / np.arange(1, n_keys + 1, dtype=np.float64) ** s
        cdf = np.cumsum(w)
        cdf /= cdf[-1]
        g = np.searchsorted(cdf, rng.random(n)).astype(np.int64) + 1
    return {"g": g,
            "v_i": rng.integers(-10**9, 10**9 + 1, n, dtype=np.int64),
            "v_f": rng.standard_normal(n, dtype=np.float32)}
