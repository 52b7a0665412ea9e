"""Grouped aggregate engine: the oracle pinned against the reference's
fixtures and brute force, then the HIP engine against the oracle."""
import math
from collections import defaultdict

import numpy as np
import pytest

from mettle_amd import abi
from oracle import oracle as orc

F = abi


def brute(keys, aggs):
    groups = defaultdict(list)
    for i in range(len(keys[0])):
        groups[tuple(int(k[i]) for k in keys)].append(i)
    out = []
    for key in sorted(groups):
        rows = groups[key]
        vals = []
        for func, col in aggs:
            if func == F.AGG_COUNT:
                vals.append(len(rows))
            elif func == F.AGG_SUM:
                vals.append(sum(col[r].item() for r in rows) if col.dtype.kind in "iu"
                            else math.fsum(float(col[r]) for r in rows))
            elif func == F.AGG_MIN:
                vals.append(min(col[r].item() for r in rows))
            elif func == F.AGG_MAX:
                vals.append(max(col[r].item() for r in rows))
            else:
                s = (sum(col[r].item() for r in rows) if col.dtype.kind in "iu"
                     else math.fsum(float(col[r]) for r in rows))
                vals.append(s / len(rows))
        out.append((key, vals))
    return out


def as_py(arr, func, t):
    if func == F.AGG_SUM and t not in (F.FLOAT, F.DOUBLE):
        return [abi.hugeint_to_int(r) for r in arr]
    return [x.item() for x in arr]


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------
def test_oracle_reference_mixed_aggregates_fixture():
    # tests/extension.rs:724-777: sales(category, amount) grouped by category
    cat = np.array([0, 1, 0, 1, 0], np.int32)  # 'A' -> 0, 'B' -> 1
    amt = np.array([10, 20, 30, 40, 50], np.int32)
    aggs = [(F.AGG_SUM, amt, F.INTEGER), (F.AGG_COUNT, None, 0), (F.AGG_MIN, amt, F.INTEGER),
            (F.AGG_MAX, amt, F.INTEGER), (F.AGG_AVG, amt, F.INTEGER)]
    keys, outs = orc.group_agg([(cat, F.DICT32)], aggs)
    assert keys[:, 0].tolist() == [0, 1]
    assert [abi.hugeint_to_int(r) for r in outs[0]] == [90, 60]
    assert outs[1].tolist() == [3, 2]
    assert outs[2].tolist() == [10, 20]
    assert outs[3].tolist() == [50, 40]
    assert outs[4].tolist() == [30.0, 30.0]


def test_oracle_reference_composite_key_fixture():
    # tests/extension.rs:876-928
    region = np.array([1, 1, 1, 1, 2, 2], np.int32)
    category = np.array([10, 10, 20, 20, 10, 10], np.int32)
    amount = np.array([100.0, 40.0, 30.0, 50.0, 80.0, 20.0])
    aggs = [(F.AGG_COUNT, None, 0), (F.AGG_SUM, amount, F.DOUBLE), (F.AGG_MIN, amount, F.DOUBLE),
            (F.AGG_MAX, amount, F.DOUBLE), (F.AGG_AVG, amount, F.DOUBLE)]
    keys, outs = orc.group_agg([(region, F.INTEGER), (category, F.INTEGER)], aggs)
    got = [tuple(k) + tuple(o[i].item() for o in outs) for i, k in enumerate(keys.tolist())]
    assert got == [(1, 10, 2, 140.0, 40.0, 100.0, 70.0), (1, 20, 2, 80.0, 30.0, 50.0, 40.0),
                   (2, 10, 2, 100.0, 20.0, 80.0, 50.0)]


def test_oracle_reference_single_key_two_sums_fixture():
    # tests/extension.rs:829-872: SUM(price) per key -> 300, 200, 300
    key = np.array([1, 1, 2, 2, 3], np.int32)
    price = np.array([100.0, 200.0, 50.0, 150.0, 300.0])
    keys, outs = orc.group_agg([(key, F.INTEGER)], [(F.AGG_SUM, price, F.DOUBLE)])
    assert keys[:, 0].tolist() == [1, 2, 3]
    assert outs[0].tolist() == [300.0, 200.0, 300.0]


def test_oracle_matches_brute_force_random():
    rng = np.random.default_rng(12)
    n = 3000
    k1 = rng.integers(-5, 5, n).astype(np.int64)
    k2 = rng.integers(0, 7, n).astype(np.int32)
    vi = rng.integers(-10**12, 10**12, n).astype(np.int64)
    vf = rng.standard_normal(n).astype(np.float32)
    aggs = [(F.AGG_SUM, vi, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, vi, F.BIGINT),
            (F.AGG_MAX, vi, F.BIGINT), (F.AGG_SUM, vf, F.FLOAT), (F.AGG_AVG, vf, F.FLOAT),
            (F.AGG_MIN, vf, F.FLOAT), (F.AGG_AVG, vi, F.BIGINT)]
    keys, outs = orc.group_agg([(k1, F.BIGINT), (k2, F.INTEGER)], aggs)
    ref = brute([k1, k2], [(f, c) for f, c, _ in aggs])
    assert [tuple(k) for k in keys.tolist()] == [k for k, _ in ref]
    for a, (f, _, t) in enumerate(aggs):
        got = as_py(outs[a], f, t)
        exp = [v[a] for _, v in ref]
        if t == F.FLOAT or (f == F.AGG_AVG):
            assert all(math.isclose(g, e, rel_tol=1e-12, abs_tol=1e-9) for g, e in zip(got, exp))
        else:
            assert got == exp


def test_oracle_hugeint_sum_exceeds_int64():
    key = np.zeros(4, np.int32)
    v = np.full(4, np.iinfo(np.int64).max, np.int64)
    _, outs = orc.group_agg([(key, F.INTEGER)], [(F.AGG_SUM, v, F.BIGINT)])
    assert abi.hugeint_to_int(outs[0][0]) == 4 * (2**63 - 1)


# ---------------------------------------------------------------------------
# GPU parity
# ---------------------------------------------------------------------------
def _gpu_ops():
    pytest.importorskip("torch")
    from mettle_amd import ops
    return ops


# FLOAT/DOUBLE SUM and AVG tolerance: the device adds in double with
# atomics, so the result depends on the order of the additions.  A sum is
# accepted when |got - exact| <= FLOAT_REL * sum(|v|) over the group's rows
# (AVG: both sides divided by the count), against the oracle's compensated
# double sum.  Measuring against sum(|v|) keeps cancellation from inflating
# the relative error of a sum near zero.
FLOAT_REL = 1e-6


def float_abs_sums(keys, col, sel=None):
    """sum(|v|) per group, in the oracle's group order."""
    _, outs = orc.group_agg(keys, [(F.AGG_SUM, np.abs(col.astype(np.float64)), F.DOUBLE),
                                   (F.AGG_COUNT, None, 0)], sel=sel)
    return outs[0], outs[1]


def assert_float_agg(got, exp, abs_sums, counts, func):
    for i, (x, y) in enumerate(zip(got, exp)):
        bound = FLOAT_REL * abs_sums[i] / (counts[i] if func == F.AGG_AVG else 1)
        assert abs(x - y) <= bound + 1e-300, (i, x, y, bound)


def compare(gk, go, rk, ro, aggs, keys=None, sel=None):
    """keys: the (np column, type) list the oracle grouped by, needed for the
    tolerance of FLOAT/DOUBLE sums."""
    assert gk.shape == rk.shape
    np.testing.assert_array_equal(gk, rk)
    for a, (f, c, t) in enumerate(aggs):
        g, r = as_py(go[a], f, t), as_py(ro[a], f, t)
        if t in (F.FLOAT, F.DOUBLE) and f in (F.AGG_SUM, F.AGG_AVG):
            abs_sums, counts = float_abs_sums(keys, c, sel)
            assert_float_agg(g, r, abs_sums, counts, f)
        else:
            assert g == r


@pytest.mark.gpu
@pytest.mark.parametrize("domain,n", [(7, 400_000), (1000, 400_000), (70_000, 400_000),
                                      (5_000_000, 400_000), (70_000, 2_000_000),
                                      (5_000_000, 3_000_000)])
def test_gpu_group_agg_parity(domain, n):
    """Below 2^20 rows every cache miss updates the global table; from 2^20
    rows on, misses of a large dense domain are spilled, sorted and reduced."""
    ops = _gpu_ops()
    rng = np.random.default_rng(domain + n)
    g = rng.integers(0, domain, n).astype(np.int64) * 3 - 17
    vi = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    vf = rng.standard_normal(n).astype(np.float32)
    aggs = [(F.AGG_SUM, vi, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, vi, F.BIGINT),
            (F.AGG_MAX, vi, F.BIGINT), (F.AGG_SUM, vf, F.FLOAT), (F.AGG_AVG, vf, F.FLOAT)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    dg, dvi, dvf = ops.to_device(g), ops.to_device(vi), ops.to_device(vf)
    dev = {id(vi): dvi, id(vf): dvf}
    groups = ops.group_agg([(dg, F.BIGINT, int(g.min()), int(g.max()))],
                           [(f, None if c is None else dev[id(c)], t) for f, c, t in aggs], n)
    gk, go = groups.fetch()
    compare(gk, go, rk, ro, aggs, keys=[(g, F.BIGINT)])


@pytest.mark.gpu
@pytest.mark.parametrize("domain,n", [(5, 300_000), (3_000_000, 300_000),
                                      (3_000_000, 1_500_000)])
def test_gpu_group_agg_int64_extreme_sums(domain, n):
    """Values over the whole int64 range (both halves of each value in use,
    low halves at the sign boundary): the 128-bit SUM, AVG, MIN and MAX are
    exact, in LDS-cached groups (5 keys) and in the global table (3M keys)."""
    ops = _gpu_ops()
    rng = np.random.default_rng(7 + domain + n)
    g = rng.integers(0, domain, n).astype(np.int64)
    special = np.array([np.iinfo(np.int64).max, np.iinfo(np.int64).min, -1, 0x7FFFFFFF,
                        -0x80000000, 0xFFFFFFFF, 0x80000000, -0x80000001,
                        np.iinfo(np.int64).max - 0x7FFFFFFF], dtype=np.int64)
    vi = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n, dtype=np.int64,
                      endpoint=True)
    m = rng.random(n) < 0.2
    vi[m] = special[rng.integers(0, len(special), int(m.sum()))]
    aggs = [(F.AGG_SUM, vi, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, vi, F.BIGINT),
            (F.AGG_MAX, vi, F.BIGINT), (F.AGG_AVG, vi, F.BIGINT)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    dg, dvi = ops.to_device(g), ops.to_device(vi)
    groups = ops.group_agg([(dg, F.BIGINT, int(g.min()), int(g.max()))],
                           [(f, None if c is None else dvi, t) for f, c, t in aggs], n)
    gk, go = groups.fetch()
    compare(gk, go, rk, ro, aggs)


@pytest.mark.gpu
@pytest.mark.parametrize("domain,n,wide", [(800, 3_000_000, False), (37, 1_200_000, False),
                                           (300, 2_000_000, True)])
def test_gpu_group_agg_staged_caches(domain, n, wide):
    """Inputs large enough for 64 or more blocks over a directly indexed
    domain: the blocks' caches are staged and reduced per slot (state words
    of up to 8 and of more than 8 words), against the oracle, with values
    over the whole int64 range, NaN-free doubles of both signs and MIN/MAX of
    every kind."""
    ops = _gpu_ops()
    rng = np.random.default_rng(domain + n + wide)
    g = rng.integers(0, domain, n).astype(np.int64) - 5
    vi = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n, dtype=np.int64,
                      endpoint=True)
    vd = rng.integers(-1000, 1000, n).astype(np.float64) * 0.5  # exact double sums
    vs = rng.integers(-100, 100, n).astype(np.int32)
    aggs = [(F.AGG_SUM, vi, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, vi, F.BIGINT),
            (F.AGG_MAX, vd, F.DOUBLE), (F.AGG_SUM, vd, F.DOUBLE), (F.AGG_MIN, vd, F.DOUBLE)]
    if wide:  # 1 + 2 + 1 + 1 + 1 + 1 + 2 + 2 state words
        aggs += [(F.AGG_SUM, vs, F.INTEGER), (F.AGG_AVG, vs, F.INTEGER)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    dev = {id(vi): ops.to_device(vi), id(vd): ops.to_device(vd), id(vs): ops.to_device(vs)}
    dg = ops.to_device(g)
    groups, names = _kernels_of(lambda: ops.group_agg(
        [(dg, F.BIGINT, int(g.min()), int(g.max()))],
        [(f, None if c is None else dev[id(c)], t) for f, c, t in aggs], n))
    assert "k_dense_reduce" in names
    gk, go = groups.fetch()
    compare(gk, go, rk, ro, aggs, keys=[(g, F.BIGINT)])


@pytest.mark.gpu
def test_gpu_group_agg_zipf_spill_with_selection():
    """A skewed key over a large dense domain (hot keys in the LDS caches,
    the tail spilled), through an input selection, with FLOAT MIN/MAX."""
    ops = _gpu_ops()
    rng = np.random.default_rng(2024)
    n = 3_000_000
    g = (np.minimum(rng.zipf(1.2, n), 2_000_000) + 5).astype(np.int64)
    vi = rng.integers(-10**12, 10**12, n).astype(np.int64)
    vf = rng.standard_normal(n).astype(np.float32)
    sel = np.flatnonzero(rng.random(n) < 0.7).astype(np.uint32)
    aggs = [(F.AGG_SUM, vi, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, vf, F.FLOAT),
            (F.AGG_MAX, vf, F.FLOAT), (F.AGG_AVG, vi, F.BIGINT), (F.AGG_SUM, vf, F.FLOAT)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs, sel=sel)
    dg, dvi, dvf = ops.to_device(g), ops.to_device(vi), ops.to_device(vf)
    dev = {id(vi): dvi, id(vf): dvf}
    dsel = ops.to_device(sel)
    groups = ops.group_agg([(dg, F.BIGINT, int(g.min()), int(g.max()))],
                           [(f, None if c is None else dev[id(c)], t) for f, c, t in aggs], n,
                           sel=dsel)
    gk, go = groups.fetch()
    compare(gk, go, rk, ro, aggs, keys=[(g, F.BIGINT)], sel=sel)


@pytest.mark.gpu
def test_gpu_group_agg_composite_keys_spill():
    """Two packed keys whose domain (about 1.4M groups) exceeds the LDS cache,
    2M rows: the spill path with composite keys, DOUBLE MIN/MAX and AVG."""
    ops = _gpu_ops()
    rng = np.random.default_rng(77)
    n = 2_000_000
    k1 = rng.integers(-300, 400, n).astype(np.int32)
    k2 = rng.integers(18000, 20000, n).astype(np.int32)  # DATE days
    vd = rng.standard_normal(n) * 1e6
    vi = rng.integers(-2**40, 2**40, n).astype(np.int64)
    aggs = [(F.AGG_COUNT, None, 0), (F.AGG_MIN, vd, F.DOUBLE), (F.AGG_MAX, vd, F.DOUBLE),
            (F.AGG_AVG, vd, F.DOUBLE), (F.AGG_SUM, vi, F.BIGINT)]
    rk, ro = orc.group_agg([(k1, F.INTEGER), (k2, F.DATE)], aggs)
    dev = {id(vd): ops.to_device(vd), id(vi): ops.to_device(vi)}
    groups = ops.group_agg([(ops.to_device(k1), F.INTEGER, int(k1.min()), int(k1.max())),
                            (ops.to_device(k2), F.DATE, int(k2.min()), int(k2.max()))],
                           [(f, None if c is None else dev[id(c)], t) for f, c, t in aggs], n)
    gk, go = groups.fetch()
    compare(gk, go, rk, ro, aggs, keys=[(k1, F.INTEGER), (k2, F.DATE)])


@pytest.mark.gpu
def test_gpu_group_agg_raw_keys_large():
    """Full-range BIGINT keys (hash table, no spill) over 1.5M rows with a
    skewed head: the hashed LDS cache with its doorkeeper, misses into the
    global hash table."""
    ops = _gpu_ops()
    rng = np.random.default_rng(101)
    n = 1_500_000
    pool = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, 200_000, dtype=np.int64)
    g = pool[np.minimum(rng.zipf(1.15, n) - 1, len(pool) - 1)]
    v = rng.integers(-10**6, 10**6, n).astype(np.int64)
    aggs = [(F.AGG_SUM, v, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, v, F.BIGINT)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    dv = ops.to_device(v)
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, np.iinfo(np.int64).min,
                             np.iinfo(np.int64).max)],
                           [(f, None if c is None else dv, t) for f, c, t in aggs], n)
    gk, go = groups.fetch()
    order = np.lexsort([gk[:, 0]])
    gk = gk[order]
    go = [o[order] for o in go]
    compare(gk, go, rk, ro, aggs)


@pytest.mark.gpu
def test_gpu_group_agg_raw_bigint_keys_and_skew():
    ops = _gpu_ops()
    rng = np.random.default_rng(99)
    n = 300_000
    # full-range BIGINT keys (no packing: raw hash path) with a hot head
    pool = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, 5000, dtype=np.int64)
    pool[0] = -1  # the all-ones pattern
    pool[1] = np.iinfo(np.int64).min
    idx = np.minimum(rng.zipf(1.3, n) - 1, len(pool) - 1)
    g = pool[idx]
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    aggs = [(F.AGG_SUM, v, F.INTEGER), (F.AGG_COUNT, None, 0), (F.AGG_MAX, v, F.INTEGER)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    dg, dv = ops.to_device(g), ops.to_device(v)
    groups = ops.group_agg([(dg, F.BIGINT, np.iinfo(np.int64).min, np.iinfo(np.int64).max)],
                           [(f, None if c is None else dv, t) for f, c, t in aggs], n)
    gk, go = groups.fetch()
    order = np.lexsort([gk[:, 0]])
    gk = gk[order]
    go = [o[order] for o in go]
    compare(gk, go, rk, ro, aggs)


@pytest.mark.gpu
def test_gpu_group_agg_composite_keys_fixture():
    ops = _gpu_ops()
    region = np.array([1, 1, 1, 1, 2, 2], np.int32)
    category = np.array([10, 10, 20, 20, 10, 10], np.int32)
    amount = np.array([100.0, 40.0, 30.0, 50.0, 80.0, 20.0])
    d = ops.to_device(amount)
    aggs = [(F.AGG_COUNT, None, 0), (F.AGG_SUM, d, F.DOUBLE), (F.AGG_MIN, d, F.DOUBLE),
            (F.AGG_MAX, d, F.DOUBLE), (F.AGG_AVG, d, F.DOUBLE)]
    groups = ops.group_agg([(ops.to_device(region), F.INTEGER, 1, 2),
                            (ops.to_device(category), F.INTEGER, 10, 20)], aggs, 6)
    keys, outs = groups.fetch()
    got = [tuple(k) + tuple(o[i].item() for o in outs) for i, k in enumerate(keys.tolist())]
    assert got == [(1, 10, 2, 140.0, 40.0, 100.0, 70.0), (1, 20, 2, 80.0, 30.0, 50.0, 40.0),
                   (2, 10, 2, 100.0, 20.0, 80.0, 50.0)]


@pytest.mark.gpu
@pytest.mark.parametrize("k", [1, 10, 100, 512])
def test_gpu_group_topk_parity(k):
    ops = _gpu_ops()
    rng = np.random.default_rng(k)
    n = 500_000
    g = rng.integers(0, 20_000, n).astype(np.int64)
    v = rng.integers(-10**9, 10**9, n).astype(np.int64)
    v[rng.integers(0, n, 2000)] = 7  # many equal sums among small groups
    aggs = [(F.AGG_SUM, v, F.BIGINT), (F.AGG_COUNT, None, 0)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    sums = np.array([abi.hugeint_to_int(r) for r in ro[0]], dtype=object)
    order = sorted(range(len(rk)), key=lambda i: (-sums[i], rk[i, 0]))[:k]
    dg, dv = ops.to_device(g), ops.to_device(v)
    groups = ops.group_agg([(dg, F.BIGINT, 0, 19_999)],
                           [(F.AGG_SUM, dv, F.BIGINT), (F.AGG_COUNT, None, 0)], n)
    tk, to = groups.topk(0, k, desc=True)
    assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order]
    assert [abi.hugeint_to_int(r) for r in to[0]] == [sums[i] for i in order]
    assert to[1].tolist() == [int(ro[1][i]) for i in order]


@pytest.mark.gpu
@pytest.mark.parametrize("k", [10, 64, 65, 100, 512])
def test_gpu_group_topk_all_ties(k):
    """Every group has the same COUNT: the order is decided by the key alone
    (ascending), whatever block or wave saw each group first."""
    ops = _gpu_ops()
    n_keys = 5000
    g = np.tile(np.arange(n_keys, dtype=np.int64)[::-1], 20)
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, 0, n_keys - 1)],
                           [(F.AGG_COUNT, None, 0)], len(g))
    tk, to = groups.topk(0, k, desc=True)
    assert tk[:, 0].tolist() == list(range(k))
    assert to[0].tolist() == [20] * k


def test_oracle_avg_decimal_divides_by_scale():
    """AVG over DECIMAL(15,2) stored as cents: DuckDB divides the hugeint sum
    by count * 10^scale (avg.cpp GetAverageDivident), so the mean of 123.45
    and 200.00 is 161.725, not 16172.5."""
    key = np.zeros(2, np.int32)
    cents = np.array([12345, 20000], np.int64)
    _, outs = orc.group_agg([(key, F.INTEGER)], [(F.AGG_AVG, cents, F.DECIMAL, 2),
                                                 (F.AGG_AVG, cents, F.BIGINT)])
    assert outs[0][0] == 161.725
    assert outs[1][0] == 16172.5


def test_oracle_avg_matches_exact_quotient():
    """The oracle's long double AVG against the exact rational mean: equal
    to within one ulp for sums within and beyond int64 (brute force with
    Python fractions)."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    n = 4000
    g = rng.integers(0, 40, n).astype(np.int32)
    v = rng.integers(-2**62, 2**62, n, dtype=np.int64)
    keys, outs = orc.group_agg([(g, F.INTEGER)], [(F.AGG_AVG, v, F.BIGINT)])
    for i, k in enumerate(keys[:, 0]):
        vals = v[g == k]
        exact = Fraction(sum(int(x) for x in vals), len(vals))
        got = outs[0][i]
        assert abs(Fraction(got) - exact) <= abs(exact) * Fraction(1, 2**52)


def _topk_expected(rk, vals, k, desc):
    """Indices of the top k groups under (value, key ascending); vals are
    exact Python numbers."""
    sign = -1 if desc else 1
    return sorted(range(len(rk)), key=lambda i: (sign * vals[i], int(rk[i, 0])))[:k]


@pytest.mark.gpu
@pytest.mark.parametrize("desc", [True, False])
def test_gpu_group_topk_avg_mixed_sign(desc):
    """ORDER BY AVG(BIGINT): sums of both signs, small negative sums (whose
    high word is -1), sums beyond +-2^64, and groups with equal averages;
    the device key must order exactly like the AVG the oracle emits."""
    ops = _gpu_ops()
    rng = np.random.default_rng(31 + desc)
    n_keys = 3000
    rows = []
    for key in range(n_keys):
        kind = key % 6
        cnt = int(rng.integers(1, 9))
        if kind == 0:  # small negative sums in (-1024, 0)
            vals = rng.integers(-300, 100, cnt)
        elif kind == 1:  # huge magnitudes: sums beyond 2^64
            vals = rng.choice([2**62 + 1, -(2**62) - 3, 2**63 - 1, -(2**63)], cnt)
        elif kind == 2:  # equal averages: -3.5 or 7.25, different counts
            vals = [-3, -4] * (cnt // 2 + 1) if key % 12 == 2 else [7, 7, 7, 8]
        else:
            vals = rng.integers(-10**12, 10**12, cnt)
        rows += [(key, int(x)) for x in vals]
    rows = np.array(rows, dtype=object)
    g = rows[:, 0].astype(np.int64)
    v = np.array([int(x) for x in rows[:, 1]], dtype=np.int64)
    perm = rng.permutation(len(g))
    g, v = g[perm], v[perm]
    aggs = [(F.AGG_AVG, v, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_SUM, v, F.BIGINT)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    avgs = [float(x) for x in ro[0]]
    dv = ops.to_device(v)
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, 0, n_keys - 1)],
                           [(F.AGG_AVG, dv, F.BIGINT), (F.AGG_COUNT, None, 0),
                            (F.AGG_SUM, dv, F.BIGINT)], len(g))
    for k in (10, 129):
        order = _topk_expected(rk, avgs, k, desc)
        tk, to = groups.topk(0, k, desc=desc)
        assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order]
        assert to[0].tolist() == [avgs[i] for i in order]
        assert [abi.hugeint_to_int(r) for r in to[2]] == \
            [abi.hugeint_to_int(ro[2][i]) for i in order]


@pytest.mark.gpu
@pytest.mark.parametrize("desc", [True, False])
def test_gpu_group_topk_avg_decimal(desc):
    """ORDER BY AVG(DECIMAL(15,2)): the key and the value divide by
    count * 100."""
    ops = _gpu_ops()
    rng = np.random.default_rng(8)
    n = 200_000
    g = rng.integers(0, 5000, n).astype(np.int64)
    cents = rng.integers(-10**9, 10**9, n).astype(np.int64)
    aggs = [(F.AGG_AVG, cents, F.DECIMAL, 2), (F.AGG_COUNT, None, 0)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    avgs = [float(x) for x in ro[0]]
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, 0, 4999)],
                           [(F.AGG_AVG, ops.to_device(cents), F.DECIMAL, 2),
                            (F.AGG_COUNT, None, 0)], n)
    order = _topk_expected(rk, avgs, 64, desc)
    tk, to = groups.topk(0, 64, desc=desc)
    assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order]
    assert to[0].tolist() == [avgs[i] for i in order]
    gk, go = groups.fetch()
    np.testing.assert_array_equal(gk, rk)
    assert go[0].tolist() == avgs


@pytest.mark.gpu
@pytest.mark.parametrize("desc", [True, False])
def test_gpu_group_topk_sums_straddling_int64(desc):
    """Sums just beyond INT64_MAX and INT64_MIN share the saturated first key
    word; the exact high and low words must order them."""
    ops = _gpu_ops()
    rng = np.random.default_rng(3)
    imax, imin = np.iinfo(np.int64).max, np.iinfo(np.int64).min
    g, v = [], []
    for j in range(1, 61):  # sums imax + j and imin - j
        g += [j, j, 100 + j, 100 + j]
        v += [imax, j, imin, -j]
    for key in range(200, 2200):  # ordinary groups in between
        g.append(key)
        v.append(int(rng.integers(-10**15, 10**15)))
    g = np.array(g, np.int64)
    v = np.array(v, np.int64)
    perm = rng.permutation(len(g))
    g, v = g[perm], v[perm]
    rk, ro = orc.group_agg([(g, F.BIGINT)], [(F.AGG_SUM, v, F.BIGINT)])
    sums = [abi.hugeint_to_int(r) for r in ro[0]]
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, int(g.min()), int(g.max()))],
                           [(F.AGG_SUM, ops.to_device(v), F.BIGINT)], len(g))
    for k in (20, 65):
        order = _topk_expected(rk, sums, k, desc)
        tk, to = groups.topk(0, k, desc=desc)
        assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order]
        assert [abi.hugeint_to_int(r) for r in to[0]] == [sums[i] for i in order]


@pytest.mark.gpu
@pytest.mark.parametrize("func,vtype,desc", [(F.AGG_MIN, F.BIGINT, True),
                                             (F.AGG_MIN, F.BIGINT, False),
                                             (F.AGG_MAX, F.INTEGER, False),
                                             (F.AGG_SUM, F.DOUBLE, True),
                                             (F.AGG_SUM, F.DOUBLE, False),
                                             (F.AGG_MAX, F.DOUBLE, True),
                                             (F.AGG_AVG, F.DOUBLE, False),
                                             (F.AGG_COUNT, None, False),
                                             (F.AGG_SUM, F.BIGINT, False)])
@pytest.mark.parametrize("k", [64, 65, 128, 129])
def test_gpu_group_topk_orders_and_k(func, vtype, desc, k):
    """Top-K by MIN, MAX, COUNT, integer and DOUBLE sums and averages, both
    directions, at K around the one- and two-row wave-list boundaries.
    DOUBLE values are small integers, so every double sum is exact in any
    order and the expected order is exact too."""
    ops = _gpu_ops()
    rng = np.random.default_rng(k * 7 + func)
    n, n_keys = 300_000, 12_000
    g = rng.integers(0, n_keys, n).astype(np.int64)
    if vtype == F.DOUBLE:
        col = rng.integers(-1000, 1000, n).astype(np.float64)
    elif vtype == F.INTEGER:
        col = rng.integers(-2**31, 2**31, n).astype(np.int32)
    else:
        col = rng.integers(-10**12, 10**12, n).astype(np.int64)
    aggs = [(func, None if func == F.AGG_COUNT else col, vtype or 0), (F.AGG_COUNT, None, 0)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    if func == F.AGG_SUM and vtype != F.DOUBLE:
        vals = [abi.hugeint_to_int(r) for r in ro[0]]
    else:
        vals = [x.item() for x in ro[0]]
    dcol = None if func == F.AGG_COUNT else ops.to_device(col)
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, 0, n_keys - 1)],
                           [(func, dcol, vtype or 0), (F.AGG_COUNT, None, 0)], n)
    order = _topk_expected(rk, vals, k, desc)
    tk, to = groups.topk(0, k, desc=desc)
    assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order]
    got = ([abi.hugeint_to_int(r) for r in to[0]] if func == F.AGG_SUM and vtype != F.DOUBLE
           else [x.item() for x in to[0]])
    assert got == [vals[i] for i in order]


@pytest.mark.gpu
def test_gpu_group_topk_many_rounds_per_block():
    """6M groups: each top-K block walks about 23K entries in several rounds
    after its seeded bound."""
    ops = _gpu_ops()
    rng = np.random.default_rng(606)
    n, n_keys = 6_500_000, 6_000_000
    g = rng.integers(0, n_keys, n).astype(np.int64)
    v = rng.integers(-10**9, 10**9, n).astype(np.int64)
    rk, ro = orc.group_agg([(g, F.BIGINT)], [(F.AGG_SUM, v, F.BIGINT)])
    sums = np.array([abi.hugeint_to_int(r) for r in ro[0]], dtype=np.int64)
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, 0, n_keys - 1)],
                           [(F.AGG_SUM, ops.to_device(v), F.BIGINT)], n)
    for desc in (True, False):
        order = np.lexsort((rk[:, 0], -sums if desc else sums))[:100]
        tk, to = groups.topk(0, 100, desc=desc)
        assert tk[:, 0].tolist() == rk[order, 0].tolist()
        assert [abi.hugeint_to_int(r) for r in to[0]] == sums[order].tolist()


def _kernels_of(fn):
    """fn()'s result and the names of the library kernels it launched."""
    import ctypes as C
    L = abi.lib()
    L.mtl_profile_reset()
    L.mtl_profile_enable(1)
    try:
        out = fn()
        abi.check(L.mtl_stream_sync(None))
    finally:
        L.mtl_profile_enable(0)
    names, i = [], 0
    name, ms, cnt = C.create_string_buffer(128), C.c_double(), C.c_uint64()
    while L.mtl_profile_get(i, name, 128, C.byref(ms), C.byref(cnt)) == 0:
        names.append(name.value.decode())
        i += 1
    L.mtl_profile_reset()
    return out, names


@pytest.mark.gpu
@pytest.mark.parametrize("func,vtype", [(F.AGG_SUM, F.BIGINT), (F.AGG_COUNT, None),
                                        (F.AGG_AVG, F.BIGINT), (F.AGG_MIN, F.INTEGER),
                                        (F.AGG_SUM, F.DOUBLE)])
@pytest.mark.parametrize("domain,holes", [(1000, False), (2048, True), (37, False)])
def test_gpu_group_topk_small_table(func, vtype, domain, holes):
    """Tables of at most 2048 entries take the one-launch top-K: every K
    (below, at and above the number of non-empty groups), both directions,
    equal values decided by the key, empty entries of a sparse key domain
    skipped, against the oracle."""
    ops = _gpu_ops()
    rng = np.random.default_rng(domain + 10 * func)
    n = 60_000
    keys = np.arange(domain, dtype=np.int64)
    if holes:
        keys = keys[rng.random(domain) < 0.6]  # about 40% of the domain never occurs
    g = rng.choice(keys, n).astype(np.int64)
    if vtype == F.DOUBLE:
        col = rng.integers(-50, 50, n).astype(np.float64)  # exact sums in any order
    elif vtype == F.INTEGER:
        col = rng.integers(-5, 5, n).astype(np.int32)  # many equal minima
    else:
        col = rng.integers(-10**12, 10**12, n).astype(np.int64)
        col[g % 7 == 0] = 3  # equal sums and averages for some groups
    aggs = [(func, None if func == F.AGG_COUNT else col, vtype or 0), (F.AGG_COUNT, None, 0)]
    rk, ro = orc.group_agg([(g, F.BIGINT)], aggs)
    if func == F.AGG_SUM and vtype != F.DOUBLE:
        vals = [abi.hugeint_to_int(r) for r in ro[0]]
    else:
        vals = [x.item() for x in ro[0]]
    dcol = None if func == F.AGG_COUNT else ops.to_device(col)
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, 0, domain - 1)],
                           [(func, dcol, vtype or 0), (F.AGG_COUNT, None, 0)], n)
    for k in sorted({1, 10, min(len(rk), 512), min(len(rk) + 5, 512), 512}):
        for desc in (True, False):
            order = _topk_expected(rk, vals, k, desc)
            (tk, to), names = _kernels_of(lambda: groups.topk(0, k, desc=desc))
            assert "k_group_topk_small" in names and "k_group_topk" not in names
            assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order], (k, desc)
            got = ([abi.hugeint_to_int(r) for r in to[0]]
                   if func == F.AGG_SUM and vtype != F.DOUBLE else [x.item() for x in to[0]])
            assert got == [vals[i] for i in order]
            assert to[1].tolist() == [int(ro[1][i]) for i in order]


@pytest.mark.gpu
def test_gpu_group_topk_small_hash_table():
    """A small hashed table (BIGINT keys over most of the int64 range, 250
    rows: at most 1024 slots and the overflow entry at any tuned load
    factor) also takes the one-launch top-K."""
    ops = _gpu_ops()
    rng = np.random.default_rng(9)
    uniq = rng.integers(-2**62, 2**62, 100, dtype=np.int64)
    g = rng.choice(uniq, 250)
    v = rng.integers(-1000, 1000, 250).astype(np.int64)
    rk, ro = orc.group_agg([(g, F.BIGINT)], [(F.AGG_SUM, v, F.BIGINT)])
    sums = [abi.hugeint_to_int(r) for r in ro[0]]
    groups = ops.group_agg([(ops.to_device(g), F.BIGINT, int(g.min()), int(g.max()))],
                           [(F.AGG_SUM, ops.to_device(v), F.BIGINT)], len(g))
    for desc in (True, False):
        order = _topk_expected(rk, sums, 100, desc)
        (tk, to), names = _kernels_of(lambda: groups.topk(0, 100, desc=desc))
        assert tk[:, 0].tolist() == [int(rk[i, 0]) for i in order]
        assert [abi.hugeint_to_int(r) for r in to[0]] == [sums[i] for i in order]
        assert "k_group_topk_small" in names
