"""Full-size parity for the scan-filter-project workload (100M rows) and the
skewed grouped aggregate (200M rows), on the same seeded generators the
benchmark uses.

Scan: every column (INTEGER, BIGINT, FLOAT, DATE, TIMESTAMP) at 1 / 10 / 50 %
selectivity, the fused filter + projection (col + k and col * k) and ORDER BY
on the filtered column, bit-exact against the C oracle (filter, projection,
overflow flag) and numpy's stable argsort (order).

Group-by: the exact per-group results of all 200M rows come from an
independent computation (pandas groupby for COUNT / SUM / MIN / MAX of the
BIGINT column and the double sum of the FLOAT column; exact integer sums also
from numpy bincounts of the 16-bit halves), checked against every group the
device emits and against its top-100; the same generator at 20M rows is
checked group by group against the C oracle.

Joins: lines 60M JOIN orders 15M through SQL on sparse BIGINT keys (the
radix-partitioned join), grouped and ungrouped, against numpy over the
generator's dense keys.
"""
import numpy as np
import pytest

from mettle_amd import abi, datagen
from oracle import oracle as orc

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N_SCAN = 100_000_000
N_GROUP = 200_000_000


@pytest.fixture(scope="module")
def ops():
    from mettle_amd import ops as _ops
    return _ops


# ---------------------------------------------------------------------------
# scan-filter-project, 100M rows
# ---------------------------------------------------------------------------
TYPES = {"c_i": abi.INTEGER, "c_b": abi.BIGINT, "c_f": abi.FLOAT, "c_d": abi.DATE,
         "c_t": abi.TIMESTAMP}


@pytest.fixture(scope="module")
def scan(ops):
    print("[full-size] generating the 100M-row scan table", flush=True)
    t = datagen.scan_table(N_SCAN)
    dev = {c: ops.to_device(v) for c, v in t.items()}
    yield t, dev
    del dev
    torch.cuda.empty_cache()


def _literal(col: str, frac: float):
    """`col < literal` passes about frac of the rows (the generator's bounds)."""
    if col == "c_i":
        return -(1 << 31) + int(frac * (1 << 32))
    if col == "c_b":
        return -(1 << 63) + int(frac * (1 << 64))
    if col == "c_f":
        from scipy.stats import norm
        return float(norm.ppf(frac))
    if col == "c_d":
        return datagen.DAYS_2000_01_01 + int(frac * (datagen.DAYS_2030_12_31 + 1 -
                                                       datagen.DAYS_2000_01_01))
    lo, hi = 946684800 * 10**6, 1924991999 * 10**6 + 1
    return lo + int(frac * (hi - lo))


# projections: (op, literal) pairs valid for the type (DATE and TIMESTAMP
# take + and - only)
PROJ = {"c_i": [(abi.PROJ_ADD, 7), (abi.PROJ_MUL, -1)],
        "c_b": [(abi.PROJ_ADD, 7), (abi.PROJ_MUL, -1)],
        "c_f": [(abi.PROJ_ADD, 0.5), (abi.PROJ_MUL, 2.5)],
        "c_d": [(abi.PROJ_ADD, 7), (abi.PROJ_SUB, 30)],
        "c_t": [(abi.PROJ_ADD, 7), (abi.PROJ_SUB, 86_400_000_000)]}


@pytest.mark.parametrize("frac", [0.01, 0.10, 0.50])
@pytest.mark.parametrize("col", list(TYPES))
def test_scan_filter_project_100m(ops, scan, col, frac):
    host, dev = scan
    t = TYPES[col]
    lit = _literal(col, frac)
    print(f"[full-size] scan {col} < {lit} ({frac:.0%})", flush=True)
    exp_sel = orc.filter([(host[col], t, abi.CMP_LT, lit)], N_SCAN)
    assert abs(len(exp_sel) / N_SCAN - frac) < 0.01
    for op, k in PROJ[col]:
        sel, vals, ovf = ops.filter_project([(dev[col], t, abi.CMP_LT, lit)], N_SCAN,
                                            dev[col], t, op, k)
        got_sel = ops.to_host(sel).view(np.uint32)
        np.testing.assert_array_equal(got_sel, exp_sel)
        exp_vals, exp_ovf = orc.project(host[col], t, exp_sel, op, k)
        assert ovf == exp_ovf
        if not exp_ovf:
            got = ops.to_host(vals)
            if t == abi.FLOAT:  # bit-exact, NaN-safe
                np.testing.assert_array_equal(got.view(np.uint32), exp_vals.view(np.uint32))
            else:
                np.testing.assert_array_equal(got, exp_vals)
    if frac == 0.10 or (col == "c_i" and frac == 0.50):
        # ORDER BY the filtered column over the selection, stable
        sel, _ = ops.filter([(dev[col], t, abi.CMP_LT, lit)], N_SCAN)
        for desc in (False, True):
            rows, n = ops.sort(dev[col], t, N_SCAN, sel=sel, desc=desc)
            assert n == len(exp_sel)
            v = host[col][exp_sel]
            if desc:  # an order-reversing map without overflow; ties keep row order
                v = -v.astype(np.float64) if t == abi.FLOAT else np.invert(v)
            order = np.argsort(v, kind="stable")
            np.testing.assert_array_equal(ops.to_host(rows).view(np.uint32), exp_sel[order])


def test_scan_equality_on_existing_date_100m(ops, scan):
    """`=` on a DATE value that occurs (about 1 / 11,323 of the rows)."""
    host, dev = scan
    lit = int(host["c_d"][12345])
    exp = orc.filter([(host["c_d"], abi.DATE, abi.CMP_EQ, lit)], N_SCAN)
    sel, n = ops.filter([(dev["c_d"], abi.DATE, abi.CMP_EQ, lit)], N_SCAN)
    np.testing.assert_array_equal(ops.to_host(sel).view(np.uint32), exp)
    assert 12345 in set(exp.tolist())


# ---------------------------------------------------------------------------
# grouped aggregate, 200M rows
# ---------------------------------------------------------------------------
FLOAT_REL = 1e-6  # FLOAT sums: |got - exact| <= FLOAT_REL * sum(|v|) per group


def _exact_groups(g, v_i, v_f):
    """Per-group COUNT, exact SUM / MIN / MAX of v_i, double SUM and sum(|v|)
    of v_f, ascending by key, computed without the library."""
    import pandas as pd
    df = pd.DataFrame({"g": g, "v": v_i, "f": v_f.astype(np.float64),
                       "a": np.abs(v_f.astype(np.float64))})
    r = df.groupby("g", sort=True).agg(c=("v", "size"), s=("v", "sum"), mn=("v", "min"),
                                       mx=("v", "max"), fs=("f", "sum"), fa=("a", "sum"))
    keys = r.index.to_numpy(np.int64)
    # the integer sums again, from exact float64 bincounts of 16-bit halves
    # (each partial sum stays below 2^53)
    lo = np.bincount(g, weights=(v_i & 0xFFFF).astype(np.float64))
    hi = np.bincount(g, weights=(v_i >> 16).astype(np.float64))
    s2 = hi[keys].astype(np.int64) * 65536 + lo[keys].astype(np.int64)
    s = r["s"].to_numpy(np.int64)
    np.testing.assert_array_equal(s, s2)
    return keys, r["c"].to_numpy(np.int64), s, r["mn"].to_numpy(np.int64), \
        r["mx"].to_numpy(np.int64), r["fs"].to_numpy(), r["fa"].to_numpy()


@pytest.mark.parametrize("dense", [False, True], ids=["zipf1.1-10M-keys", "1000-keys"])
def test_group_agg_200m(ops, dense):
    print(f"[full-size] generating the 200M-row group table (dense={dense})", flush=True)
    t = datagen.group_table(N_GROUP, dense_keys=1000 if dense else None)
    g, v_i, v_f = t["g"], t["v_i"], t["v_f"]
    gmax = int(g.max())
    dg, dvi, dvf = ops.to_device(g), ops.to_device(v_i), ops.to_device(v_f)
    aggs = [(abi.AGG_SUM, dvi, abi.BIGINT), (abi.AGG_COUNT, None, 0),
            (abi.AGG_MIN, dvi, abi.BIGINT), (abi.AGG_MAX, dvi, abi.BIGINT),
            (abi.AGG_SUM, dvf, abi.FLOAT), (abi.AGG_AVG, dvf, abi.FLOAT)]
    print("[full-size] device group-by", flush=True)
    groups = ops.group_agg([(dg, abi.BIGINT, 0, gmax)], aggs, N_GROUP)
    gk, go = groups.fetch()
    top_k, top_o = groups.topk(0, 100, desc=True)
    groups.release()
    del dg, dvi, dvf
    torch.cuda.empty_cache()
    # size-independent properties first
    lower = go[0]["lower"].view(np.int64)
    assert np.all(go[0]["upper"] == (lower >> 63))  # every sum fits int64 here
    gsum = lower
    assert int(go[1].sum()) == N_GROUP
    assert int(gsum.sum()) == int(v_i.sum())
    assert int(go[2].min()) == int(v_i.min()) and int(go[3].max()) == int(v_i.max())
    print("[full-size] exact per-group results on the host", flush=True)
    keys, cnt, s, mn, mx, fs, fa = _exact_groups(g, v_i, v_f)
    np.testing.assert_array_equal(gk[:, 0], keys)
    np.testing.assert_array_equal(go[1], cnt)
    np.testing.assert_array_equal(gsum, s)
    np.testing.assert_array_equal(go[2], mn)
    np.testing.assert_array_equal(go[3], mx)
    assert np.all(np.abs(go[4] - fs) <= FLOAT_REL * fa + 1e-300)
    assert np.all(np.abs(go[5] - fs / cnt) <= FLOAT_REL * fa / cnt + 1e-300)
    # top-100 by (SUM(v_i) DESC, g ASC)
    order = np.lexsort((keys, -s))[:100]
    assert top_k[:, 0].tolist() == keys[order].tolist()
    assert [abi.hugeint_to_int(r) for r in top_o[0]] == s[order].tolist()
    np.testing.assert_array_equal(top_o[1], cnt[order])
    np.testing.assert_array_equal(top_o[2], mn[order])
    np.testing.assert_array_equal(top_o[3], mx[order])


def test_group_agg_20m_matches_oracle(ops):
    """The config-4 generator at 20M rows, group by group against the oracle
    (integers bit-exact; FLOAT sums at 1e-6 of sum(|v|))."""
    t = datagen.group_table(20_000_000)
    g, v_i, v_f = t["g"], t["v_i"], t["v_f"]
    aggs = [(abi.AGG_SUM, v_i, abi.BIGINT), (abi.AGG_COUNT, None, 0),
            (abi.AGG_MIN, v_i, abi.BIGINT), (abi.AGG_MAX, v_i, abi.BIGINT),
            (abi.AGG_SUM, v_f, abi.FLOAT), (abi.AGG_AVG, v_f, abi.FLOAT),
            (abi.AGG_AVG, v_i, abi.BIGINT)]
    rk, ro = orc.group_agg([(g, abi.BIGINT)], aggs)
    fa = np.bincount(g, weights=np.abs(v_f.astype(np.float64)))[rk[:, 0]]
    dev = {id(v_i): ops.to_device(v_i), id(v_f): ops.to_device(v_f)}
    groups = ops.group_agg([(ops.to_device(g), abi.BIGINT, 0, int(g.max()))],
                           [(f, None if c is None else dev[id(c)], ty) for f, c, ty in aggs],
                           len(g))
    gk, go = groups.fetch()
    np.testing.assert_array_equal(gk, rk)
    assert [abi.hugeint_to_int(r) for r in go[0]] == [abi.hugeint_to_int(r) for r in ro[0]]
    for a in (1, 2, 3, 6):
        np.testing.assert_array_equal(go[a], ro[a])
    cnt = ro[1]
    assert np.all(np.abs(go[4] - ro[4]) <= FLOAT_REL * fa + 1e-300)
    assert np.all(np.abs(go[5] - ro[5]) <= FLOAT_REL * fa / cnt + 1e-300)


# ---------------------------------------------------------------------------
# joins at the star workload's sizes (lines 60M, orders 15M) through SQL
# ---------------------------------------------------------------------------
def test_sql_join_grouped_at_full_size_sparse_keys(ops):
    """lines 60M JOIN orders 15M on sparse BIGINT keys (the radix-partitioned
    join: a grouped consumer, keys too sparse for the dense table), filtered
    by o_date, grouped by o_prio; then the same query as a projection-free
    ungrouped aggregate.  Expected from numpy over the generator's dense
    keys (l_order indexes orders)."""
    from mettle_amd import sql as S
    print("[full-size] generating the star tables", flush=True)
    _, orders, lines = datagen.star_tables()
    mult = np.int64(2654435761)
    o_sk = orders["o_key"].astype(np.int64) * mult - (1 << 50)
    l_sk = lines["l_order"].astype(np.int64) * mult - (1 << 50)
    keep = []

    def reg(name, n, cols):
        spec = []
        for cname, t, v, uq in cols:
            d = ops.to_device(np.ascontiguousarray(v))
            keep.append(d)
            spec.append((cname, t, d.data_ptr(), int(v.min()), int(v.max()), uq))
        S.register_table(name, n, spec)

    S.drop()
    try:
        reg("orders", len(o_sk), [("o_skey", abi.BIGINT, o_sk, True),
                                  ("o_date", abi.DATE, orders["o_date"], False),
                                  ("o_prio", abi.INTEGER, orders["o_prio"], False)])
        reg("lines", len(l_sk), [("l_skey", abi.BIGINT, l_sk, False),
                                 ("price", abi.BIGINT, lines["price"], False)])
        cutoff = int(datagen.DAYS_1992_01_01 + 1170)
        import datetime
        day = (datetime.date(1970, 1, 1) + datetime.timedelta(days=cutoff)).isoformat()
        out = S.execute(f"SELECT o.o_prio, COUNT(*) AS n, SUM(l.price) AS s FROM lines l "
                        f"JOIN orders o ON l.l_skey = o.o_skey WHERE o.o_date < DATE '{day}' "
                        f"GROUP BY o.o_prio ORDER BY o.o_prio")
        oi = lines["l_order"] - 1
        m = orders["o_date"][oi] < cutoff
        prio = orders["o_prio"][oi][m]
        n_exp = np.bincount(prio, minlength=5)
        s_exp = np.bincount(prio, weights=lines["price"][m].astype(np.float64), minlength=5)
        # exact integer sums (float64 weights are exact below 2^53: 6e14 max)
        assert [int(x) for x in out["o_prio"][2]] == [p for p in range(5) if n_exp[p]]
        assert [int(x) for x in out["n"][2]] == [int(n_exp[p]) for p in range(5) if n_exp[p]]
        assert [int(x) for x in out["s"][2]] == [int(s_exp[p]) for p in range(5) if n_exp[p]]
        out = S.execute(f"SELECT COUNT(*) AS n, MAX(l.price) AS hi FROM lines l JOIN orders o "
                        f"ON l.l_skey = o.o_skey WHERE o.o_date < DATE '{day}'")
        assert int(out["n"][2][0]) == int(m.sum())
        assert int(out["hi"][2][0]) == int(lines["price"][m].max())
    finally:
        S.drop()
        keep.clear()
        torch.cuda.empty_cache()
