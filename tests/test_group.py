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


def compare(gk, go, rk, ro, aggs, float_rel=1e-9):
    assert gk.shape == rk.shape
    np.testing.assert_array_equal(gk, rk)
    for a, (f, _, t) in enumerate(aggs):
        g, r = as_py(go[a], f, t), as_py(ro[a], f, t)
        if t in (F.FLOAT, F.DOUBLE) and f in (F.AGG_SUM, F.AGG_AVG):
            for x, y in zip(g, r):
                assert math.isclose(x, y, rel_tol=float_rel, abs_tol=1e-9), (x, y)
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
    compare(gk, go, rk, ro, aggs)


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
    compare(gk, go, rk, ro, aggs)


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
    compare(gk, go, rk, ro, aggs)


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
