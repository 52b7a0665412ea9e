"""GPU parity: every operator of libmettle_hip.so against the oracle on the
same seeded inputs, bit-exact (integer/byte/index work and IEEE f32/f64
element-wise results)."""
import numpy as np
import pytest

from mettle_amd import abi, datagen
from oracle import oracle as orc
from tests import golden

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ops():
    from mettle_amd import ops as _ops
    return _ops


TYPES = [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64), (abi.DATE, np.int32),
         (abi.TIMESTAMP, np.int64), (abi.DICT8, np.uint8), (abi.FLOAT, np.float32),
         (abi.DOUBLE, np.float64)]


def make_col(rng, dtype, n):
    if np.issubdtype(dtype, np.floating):
        return rng.standard_normal(n).astype(dtype)
    info = np.iinfo(dtype)
    return rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)


@pytest.mark.parametrize("t,dtype", TYPES)
@pytest.mark.parametrize("op", [abi.CMP_EQ, abi.CMP_LT, abi.CMP_LE, abi.CMP_GT, abi.CMP_GE])
@pytest.mark.parametrize("n", [0, 1, 4095, 4096, 100_003])
def test_filter_parity(ops, t, dtype, op, n):
    rng = np.random.default_rng(1000 + t * 7 + op + n)
    col = make_col(rng, dtype, max(n, 1))[:n]
    lit = col[n // 3].item() if n else 0
    ref = orc.filter([(col, t, op, lit)], n)
    dcol = ops.to_device(col) if n else ops.to_device(np.zeros(1, dtype))
    got, cnt = ops.filter([(dcol, t, op, lit)], n)
    assert cnt == len(ref)
    np.testing.assert_array_equal(ops.to_host(got, np.uint32), ref)


@pytest.mark.parametrize("n", [0, 1, 4097, 300_001])
@pytest.mark.parametrize("seed", range(4))
def test_filter_or_runs(ops, n, seed):
    """mtl_filter_or against set algebra over the oracle's one-predicate
    filters: each OR run is the union of its predicates' row sets, the runs
    are intersected."""
    rng = np.random.default_rng(50 + seed)
    cols = {abi.INTEGER: rng.integers(-1000, 1000, max(n, 1)).astype(np.int32)[:n],
            abi.BIGINT: rng.integers(-50, 50, max(n, 1)).astype(np.int64)[:n],
            abi.DOUBLE: rng.standard_normal(max(n, 1))[:n],
            abi.DICT8: rng.integers(0, 9, max(n, 1)).astype(np.uint8)[:n]}
    dev = {t: ops.to_device(c if n else np.zeros(1, c.dtype)) for t, c in cols.items()}
    k = int(rng.integers(2, 9))
    preds, dpreds = [], []
    for i in range(k):
        t = [abi.INTEGER, abi.BIGINT, abi.DOUBLE, abi.DICT8][int(rng.integers(0, 4))]
        op = int(rng.integers(0, 5))
        c = cols[t]
        lit = c[int(rng.integers(0, n))].item() if n else 0
        preds.append((c, t, op, lit))
        dpreds.append((dev[t], t, op, lit))
    or_next = int(rng.integers(0, 1 << (k - 1)))
    keep = np.ones(n, bool)
    run = np.zeros(n, bool)
    for i, p in enumerate(preds):
        m = np.zeros(n, bool)
        m[orc.filter([p], n)] = True
        run |= m
        if not (or_next >> i) & 1:
            keep &= run
            run = np.zeros(n, bool)
    got, cnt = ops.filter_or(dpreds, or_next, n)
    assert cnt == int(keep.sum())
    np.testing.assert_array_equal(ops.to_host(got, np.uint32), np.flatnonzero(keep))


def test_filter_multi_predicate_and_refine(ops):
    rng = np.random.default_rng(2)
    n = 300_001
    a = rng.integers(-1000, 1000, n).astype(np.int32)
    b = rng.standard_normal(n).astype(np.float32)
    c = rng.integers(0, 5, n).astype(np.uint8)
    preds = [(a, abi.INTEGER, abi.CMP_GE, -10), (b, abi.FLOAT, abi.CMP_LT, 0.25),
             (c, abi.DICT8, abi.CMP_EQ, 3)]
    ref = orc.filter(preds, n)
    da, db, dc = ops.to_device(a), ops.to_device(b), ops.to_device(c)
    got, _ = ops.filter([(da, abi.INTEGER, abi.CMP_GE, -10), (db, abi.FLOAT, abi.CMP_LT, 0.25),
                         (dc, abi.DICT8, abi.CMP_EQ, 3)], n)
    np.testing.assert_array_equal(ops.to_host(got, np.uint32), ref)
    # refine an existing selection (device count)
    sel, cnt = ops.filter([(da, abi.INTEGER, abi.CMP_GE, -10)], n)
    cnt_t = torch.tensor([cnt], dtype=torch.int32, device="cuda")
    got2, _ = ops.filter([(db, abi.FLOAT, abi.CMP_LT, 0.25), (dc, abi.DICT8, abi.CMP_EQ, 3)], n,
                         in_sel=sel.contiguous(), in_count=cnt_t)
    np.testing.assert_array_equal(ops.to_host(got2, np.uint32), ref)


@pytest.mark.parametrize("lit", [2**40, -2**40, 2**31, -2**31 - 1])
def test_filter_int_column_wide_literal(ops, lit):
    v = np.array([-2**31, -5, 0, 5, 2**31 - 1] * 1000, dtype=np.int32)
    d = ops.to_device(v)
    for op in (abi.CMP_EQ, abi.CMP_LT, abi.CMP_LE, abi.CMP_GT, abi.CMP_GE):
        ref = orc.filter([(v, abi.INTEGER, op, lit)], len(v))
        got, _ = ops.filter([(d, abi.INTEGER, op, lit)], len(v))
        np.testing.assert_array_equal(ops.to_host(got, np.uint32), ref)


@pytest.mark.parametrize("t,dtype", [(abi.DICT8, np.uint8), (abi.INTEGER, np.int32),
                                     (abi.BIGINT, np.int64), (abi.DOUBLE, np.float64)])
def test_filter_unaligned_columns(ops, t, dtype):
    """Columns that start off the vector-load alignment (scalar loads)."""
    rng = np.random.default_rng(300 + t)
    col = make_col(rng, dtype, 50_001)
    lit = col[777].item()
    d = ops.to_device(col)[1:]
    ref = orc.filter([(col[1:], t, abi.CMP_LE, lit)], len(col) - 1)
    got, _ = ops.filter([(d, t, abi.CMP_LE, lit)], len(col) - 1)
    np.testing.assert_array_equal(ops.to_host(got, np.uint32), ref)


@pytest.mark.parametrize("frac", [0.0, 0.001, 0.5, 1.0])
def test_filter_many_tiles(ops, frac):
    """Thousands of 4096-row tiles (look-back windows well past 64 tiles),
    from no row to every row passing."""
    rng = np.random.default_rng(int(frac * 1000) + 7)
    n = 12_000_017
    col = rng.integers(0, 1 << 30, n, dtype=np.int32)
    lit = int(frac * (1 << 30))
    ref = orc.filter([(col, abi.INTEGER, abi.CMP_LT, lit)], n)
    got, cnt = ops.filter([(ops.to_device(col), abi.INTEGER, abi.CMP_LT, lit)], n)
    assert cnt == len(ref)
    np.testing.assert_array_equal(ops.to_host(got, np.uint32), ref)


def test_filter_float_nan_semantics(ops):
    col = np.array([np.nan, 1.0, -np.inf, np.inf, np.nan, 0.0, -0.0], dtype=np.float32)
    d = ops.to_device(col)
    for op, lit in [(abi.CMP_EQ, float("nan")), (abi.CMP_GT, float("inf")),
                    (abi.CMP_LT, float("nan")), (abi.CMP_EQ, 0.0), (abi.CMP_GE, -0.0)]:
        ref = orc.filter([(col, abi.FLOAT, op, lit)], len(col))
        got, _ = ops.filter([(d, abi.FLOAT, op, lit)], len(col))
        np.testing.assert_array_equal(ops.to_host(got, np.uint32), ref)


def test_compare_mask(ops):
    rng = np.random.default_rng(4)
    a = rng.integers(0, 100, 10_000).astype(np.int32)
    m = ops.compare([(ops.to_device(a), abi.INTEGER, abi.CMP_LT, 40)], len(a))
    np.testing.assert_array_equal(ops.to_host(m), (a < 40).astype(np.uint8))


@pytest.mark.parametrize("dtype", [np.uint8, np.int32, np.int64])
def test_gather(ops, dtype):
    rng = np.random.default_rng(5)
    col = make_col(rng, dtype, 50_000)
    sel = np.sort(rng.choice(50_000, 20_000, replace=False)).astype(np.uint32)
    got = ops.gather(ops.to_device(col), ops.to_device(sel))
    np.testing.assert_array_equal(ops.to_host(got), col[sel])


@pytest.mark.parametrize("t,dtype", [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64),
                                     (abi.DECIMAL, np.int64), (abi.FLOAT, np.float32),
                                     (abi.DOUBLE, np.float64)])
@pytest.mark.parametrize("op", [abi.PROJ_ADD, abi.PROJ_MUL, abi.PROJ_SUB, abi.PROJ_SUBFROM])
def test_project_parity(ops, t, dtype, op):
    rng = np.random.default_rng(6 + op + t)
    col = make_col(rng, dtype, 70_000)
    if not np.issubdtype(dtype, np.floating):
        col = (col // 4096).astype(dtype)  # room so most rows do not overflow
    sel = np.sort(rng.choice(70_000, 30_000, replace=False)).astype(np.uint32)
    s = 1.1 if np.issubdtype(dtype, np.floating) else 977
    ref, rovf = orc.project(col, t, sel, op, s)
    got, govf = ops.project(ops.to_device(col), t, ops.to_device(sel), op, s)
    np.testing.assert_array_equal(ops.to_host(got).view(np.uint8), ref.view(np.uint8))
    assert govf == rovf


def test_project_overflow_detected(ops):
    col = np.array([2**31 - 10, 5], dtype=np.int32)
    _, ovf = ops.project(ops.to_device(col), abi.INTEGER, None, abi.PROJ_ADD, 9)
    assert not ovf
    _, ovf = ops.project(ops.to_device(col), abi.INTEGER, None, abi.PROJ_ADD, 10)
    assert ovf


@pytest.mark.parametrize("t,dtype,op", [(abi.DATE, np.int32, abi.PROJ_ADD),
                                        (abi.DATE, np.int32, abi.PROJ_SUB),
                                        (abi.TIMESTAMP, np.int64, abi.PROJ_ADD),
                                        (abi.TIMESTAMP, np.int64, abi.PROJ_SUB)])
def test_project_date_timestamp(ops, t, dtype, op):
    rng = np.random.default_rng(60 + t + op)
    col = (make_col(rng, dtype, 20_000) // 65536).astype(dtype)
    s = 86_400_000_000 if t == abi.TIMESTAMP else 30
    ref, rovf = orc.project(col, t, None, op, s)
    got, govf = ops.project(ops.to_device(col), t, None, op, s)
    np.testing.assert_array_equal(ops.to_host(got), ref)
    assert govf == rovf


@pytest.mark.parametrize("t", [abi.DATE, abi.TIMESTAMP])
@pytest.mark.parametrize("op", [abi.PROJ_MUL, abi.PROJ_SUBFROM])
def test_project_date_timestamp_rejects_mul(ops, t, op):
    col = np.arange(8, dtype=np.int32 if t == abi.DATE else np.int64)
    with pytest.raises(abi.Unsupported):
        ops.project(ops.to_device(col), t, None, op, 2)
    with pytest.raises(abi.Unsupported):
        ops.filter_project([], len(col), ops.to_device(col), t, op, 2)


FP_TYPES = [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64), (abi.DECIMAL, np.int64),
            (abi.DATE, np.int32), (abi.TIMESTAMP, np.int64), (abi.FLOAT, np.float32),
            (abi.DOUBLE, np.float64)]


@pytest.mark.parametrize("t,dtype", FP_TYPES)
@pytest.mark.parametrize("n", [0, 1, 4095, 4097, 250_003])
def test_filter_project_parity(ops, t, dtype, n):
    """The fused entry equals mtl_filter followed by the oracle's projection."""
    rng = np.random.default_rng(70 + t * 11 + n)
    a = rng.integers(-1000, 1000, max(n, 1)).astype(np.int32)[:n]
    col = make_col(rng, dtype, max(n, 1))[:n]
    if not np.issubdtype(dtype, np.floating):
        col = (col // 4096).astype(dtype)
    fp = np.issubdtype(dtype, np.floating)
    op = abi.PROJ_ADD if t in (abi.DATE, abi.TIMESTAMP) else abi.PROJ_MUL
    s = 1.1 if fp else 3
    preds_h = [(a, abi.INTEGER, abi.CMP_LT, 250)]
    sel = orc.filter(preds_h, n)
    ref, rovf = orc.project(col, t, sel, op, s)
    da = ops.to_device(a) if n else ops.to_device(np.zeros(1, np.int32))
    dc = ops.to_device(col) if n else ops.to_device(np.zeros(1, dtype))
    gsel, gvals, govf = ops.filter_project([(da, abi.INTEGER, abi.CMP_LT, 250)], n, dc, t, op, s)
    np.testing.assert_array_equal(ops.to_host(gsel, np.uint32), sel)
    np.testing.assert_array_equal(ops.to_host(gvals).view(np.uint8), ref.view(np.uint8))
    assert govf == rovf


@pytest.mark.parametrize("t,dtype", FP_TYPES)
@pytest.mark.parametrize("n", [0, 1, 4097, 1_000_003])
@pytest.mark.parametrize("second", [False, True])
def test_filter_project_same_column(ops, t, dtype, n, second):
    """Projection of the first predicate's column: the filter pass stages the
    passing values for the expansion (optionally with a second predicate on
    another column)."""
    rng = np.random.default_rng(90 + t * 5 + n + second)
    col = make_col(rng, dtype, max(n, 1))[:n]
    fp = np.issubdtype(dtype, np.floating)
    if not fp:
        col = (col // 4096).astype(dtype)
    b = rng.integers(0, 100, max(n, 1)).astype(np.int32)[:n]
    lit = col[n // 4].item() if n else 0
    op = abi.PROJ_SUB if t in (abi.DATE, abi.TIMESTAMP) else abi.PROJ_ADD
    s_ = 2.5 if fp else 11
    preds_h = [(col, t, abi.CMP_GE, lit)] + ([(b, abi.INTEGER, abi.CMP_LT, 60)] if second else [])
    sel = orc.filter(preds_h, n)
    ref, rovf = orc.project(col, t, sel, op, s_)
    dc = ops.to_device(col) if n else ops.to_device(np.zeros(1, dtype))
    db = ops.to_device(b) if n else ops.to_device(np.zeros(1, np.int32))
    preds_d = [(dc, t, abi.CMP_GE, lit)] + ([(db, abi.INTEGER, abi.CMP_LT, 60)] if second else [])
    gsel, gvals, govf = ops.filter_project(preds_d, n, dc, t, op, s_)
    np.testing.assert_array_equal(ops.to_host(gsel, np.uint32), sel)
    np.testing.assert_array_equal(ops.to_host(gvals).view(np.uint8), ref.view(np.uint8))
    assert govf == rovf


def test_filter_project_refine_and_overflow(ops):
    """An input selection with a device count; overflow of a passing row is
    flagged, overflow of a failing row is not."""
    rng = np.random.default_rng(80)
    n = 100_000
    a = rng.integers(0, 100, n).astype(np.int32)
    v = rng.integers(-1000, 1000, n).astype(np.int32)
    big = np.where(a >= 50)[0][:3]
    v[big] = 2**31 - 5
    da, dv = ops.to_device(a), ops.to_device(v)
    sel0, cnt0 = ops.filter([(da, abi.INTEGER, abi.CMP_GE, 10)], n)
    cnt_t = torch.tensor([cnt0], dtype=torch.int32, device="cuda")
    # rows with a < 50 never carry the large values: no overflow
    s1, v1, o1 = ops.filter_project([(da, abi.INTEGER, abi.CMP_LT, 50)], n, dv, abi.INTEGER,
                                    abi.PROJ_ADD, 10, in_sel=sel0.contiguous(), in_count=cnt_t)
    ref_sel = orc.filter([(a, abi.INTEGER, abi.CMP_GE, 10), (a, abi.INTEGER, abi.CMP_LT, 50)], n)
    ref_v, ref_o = orc.project(v, abi.INTEGER, ref_sel, abi.PROJ_ADD, 10)
    np.testing.assert_array_equal(ops.to_host(s1, np.uint32), ref_sel)
    np.testing.assert_array_equal(ops.to_host(v1), ref_v)
    assert not o1 and not ref_o
    _, _, o2 = ops.filter_project([(da, abi.INTEGER, abi.CMP_GE, 50)], n, dv, abi.INTEGER,
                                  abi.PROJ_ADD, 10)
    assert o2


@pytest.mark.parametrize("strategy", [abi.JOIN_DENSE, abi.JOIN_HASH])
@pytest.mark.parametrize("t,dtype", [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64)])
def test_join_parity(ops, strategy, t, dtype):
    rng = np.random.default_rng(7)
    nb, npr = 50_000, 400_000
    base = -20_000 if t == abi.INTEGER else -(1 << 40)
    bk = (base + rng.permutation(120_000)[:nb]).astype(dtype)
    pk = (base + rng.integers(-5000, 130_000, npr)).astype(dtype)
    ref_p, ref_b = orc.join(bk, t, pk, t)
    mn, mx = int(bk.min()), int(bk.max())
    slots = mx - mn + 1 if strategy == abi.JOIN_DENSE else 0
    if strategy == abi.JOIN_DENSE and t == abi.BIGINT:
        pass  # dense over a shifted 64-bit base is still a small span
    gp, gb = ops.join(strategy, ops.to_device(bk), t, ops.to_device(pk), t, mn, slots)
    np.testing.assert_array_equal(ops.to_host(gp, np.uint32), ref_p)
    np.testing.assert_array_equal(ops.to_host(gb, np.uint32), ref_b)


def test_join_hash_extreme_keys(ops):
    # INT64_MIN (the empty sentinel) and -1 (the reference's sentinel) are keys
    bk = np.array([np.iinfo(np.int64).min, -1, 0, np.iinfo(np.int64).max], dtype=np.int64)
    pk = np.array([0, -1, 5, np.iinfo(np.int64).min, np.iinfo(np.int64).max, -1], dtype=np.int64)
    gp, gb = ops.join(abi.JOIN_HASH, ops.to_device(bk), abi.BIGINT, ops.to_device(pk),
                      abi.BIGINT)
    assert list(ops.to_host(gp, np.uint32)) == [0, 1, 3, 4, 5]
    assert list(ops.to_host(gb, np.uint32)) == [2, 1, 0, 3, 1]


def _by_probe(gp, gb):
    p = gp.astype(np.int64)
    o = np.argsort(p, kind="stable")
    return p[o], gb[o]


@pytest.mark.parametrize("t,dtype", [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64)])
@pytest.mark.parametrize("nb,npr", [(1, 1000), (2048, 100_000), (2049, 50_000),
                                    (50_000, 400_000), (3_000_000, 6_000_000)])
def test_join_radix_parity(ops, t, dtype, nb, npr):
    """LDS-staged partitions: every (probe, build) pair of the oracle's join
    (partition order, so compared by probe row); one partition at 2048 build
    rows, 2 at 2049, 2048 partitions at 3M."""
    rng = np.random.default_rng(nb + npr + t)
    span = 3 * nb + 10
    base = -(span // 2) if t == abi.INTEGER else -(1 << 40)
    bk = (base + rng.permutation(span)[:nb]).astype(dtype)
    pk = (base + rng.integers(-5, span + 5, npr)).astype(dtype)
    ref_p, ref_b = orc.join(bk, t, pk, t)
    gp, gb = ops.join(abi.JOIN_RADIX, ops.to_device(bk), t, ops.to_device(pk), t)
    p, b = _by_probe(ops.to_host(gp, np.uint32), ops.to_host(gb, np.uint32))
    np.testing.assert_array_equal(p, ref_p.astype(np.int64))
    np.testing.assert_array_equal(b, ref_b)


def test_join_radix_selections(ops):
    rng = np.random.default_rng(9)
    nb, npr = 40_000, 300_000
    bk = (rng.choice(1 << 27, nb, replace=False) - (1 << 26)).astype(np.int64)
    pk = bk[rng.integers(0, nb, npr)].copy()
    pk[rng.random(npr) < 0.3] = 12345678  # misses (not a build key)
    bk[bk == 12345678] = 12345679
    bsel = np.sort(rng.choice(nb, nb // 2, replace=False)).astype(np.uint32)
    psel = np.sort(rng.choice(npr, npr // 3, replace=False)).astype(np.uint32)
    ref_p, ref_b = orc.join(bk, abi.BIGINT, pk, abi.BIGINT, bsel, psel)
    gp, gb = ops.join(abi.JOIN_RADIX, ops.to_device(bk), abi.BIGINT, ops.to_device(pk),
                      abi.BIGINT, bsel=ops.to_device(bsel.view(np.int32)),
                      psel=ops.to_device(psel.view(np.int32)))
    p, b = _by_probe(ops.to_host(gp, np.uint32), ops.to_host(gb, np.uint32))
    np.testing.assert_array_equal(p, ref_p.astype(np.int64))
    np.testing.assert_array_equal(b, ref_b)


@pytest.mark.parametrize("nb", [4, 5000])
def test_join_radix_extreme_keys(ops, nb):
    # INT64_MIN (the empty sentinel) and -1 are keys, in one partition or many
    rng = np.random.default_rng(nb)
    extra = rng.choice(10**9, nb - 4, replace=False).astype(np.int64) + 10
    bk = np.concatenate([np.array([np.iinfo(np.int64).min, -1, 0, np.iinfo(np.int64).max],
                                  dtype=np.int64), extra])
    pk = np.array([0, -1, 5, np.iinfo(np.int64).min, np.iinfo(np.int64).max, -1], dtype=np.int64)
    gp, gb = ops.join(abi.JOIN_RADIX, ops.to_device(bk), abi.BIGINT, ops.to_device(pk),
                      abi.BIGINT)
    p, b = _by_probe(ops.to_host(gp, np.uint32), ops.to_host(gb, np.uint32))
    assert list(p) == [0, 1, 3, 4, 5]
    assert list(b) == [2, 1, 0, 3, 1]


@pytest.mark.parametrize("strategy", [abi.JOIN_DENSE, abi.JOIN_HASH, abi.JOIN_RADIX])
def test_join_duplicate_build_key_reported(ops, strategy):
    bk = np.array([3, 7, 3], dtype=np.int32)
    with pytest.raises(abi.Unsupported):
        ops.join(strategy, ops.to_device(bk), abi.INTEGER, ops.to_device(bk), abi.INTEGER, 3,
                 5 if strategy == abi.JOIN_DENSE else 0)


# ---------------------------------------------------------------------------
# star pipeline
# ---------------------------------------------------------------------------
def device_desc(ops, tables, builder):
    dev = {name: {c: ops.to_device(v) for c, v in t.items()}
           for name, t in zip(["customers", "orders", "lines"], tables)}
    d = builder(lambda t, c: dev[t][c].data_ptr())
    return d, dev


def host_desc(tables, builder):
    by = dict(zip(["customers", "orders", "lines"], tables))
    return builder(lambda t, c: by[t][c].ctypes.data)


@pytest.mark.parametrize("decimal", [True, False])
@pytest.mark.parametrize("boundary", [False, True])
def test_star_reference_q3_fixture(ops, decimal, boundary):
    tables = golden.q3_tables(decimal, boundary)
    d, _keep = device_desc(ops, tables, lambda p: golden.q3_desc(p, tables, decimal))
    rows = ops.star_topk(d)
    o = tables[1]
    got = [(int(o["o_key"][r.p_row]), r.sum_f64, int(o["o_date"][r.p_row]),
            int(o["o_prio"][r.p_row])) for r in rows]
    if decimal:
        got = [(k, s / 10_000.0, dt, pr) for k, s, dt, pr in got]
    assert got == golden.Q3_EXPECTED


def _star_parity(ops, tables, build):
    st, ref = orc.star_topk(host_desc(tables, build))
    assert st == abi.MTL_OK
    d, _keep = device_desc(ops, tables, build)
    rows = ops.star_topk(d)
    assert [(r.p_row, r.count, r.sum_i64) for r in rows] == \
        [(r.p_row, r.count, r.sum_i64) for r in ref]
    return rows


@pytest.mark.parametrize("scale,k", [(0.002, 10), (0.01, 10), (0.01, 100), (0.05, 512),
                                     (0.1, 10), (1.0, 10)])
def test_star_parity_random(ops, scale, k):
    """Scales 0.1 and 1.0 have parent tables of 1.5M and 15M rows, which take
    the persistent probe pass (dimension spans of 150K and 1.5M keys: all of
    the first and 1,179,648 keys of the second are probed in LDS)."""
    tables = datagen.star_tables(scale=scale, workload=30)
    n_c, n_o, n_l = (len(tables[0]["c_key"]), len(tables[1]["o_key"]),
                     len(tables[2]["l_order"]))
    _star_parity(ops, tables, lambda p: datagen.star_desc(p, n_c, n_o, n_l, k=k))


@pytest.mark.parametrize("shuffle_orders,shuffle_lines", [(False, False), (True, False),
                                                           (False, True)])
def test_star_parity_wide_dimension_span(ops, shuffle_orders, shuffle_lines):
    """Customer keys stretched x20 (span 3M): probes below 1,179,648 read the
    LDS copy of the bitmap, the rest read it in HBM.  Shuffled orders make
    the parent's mark keys unsorted, so few lanes share a bitmap word;
    shuffled lines make the fact scan's foreign keys unclustered, so its
    two-word probe windows miss and rows probe one by one."""
    cust, orders, lines = datagen.star_tables(scale=0.1, workload=32)
    cust["c_key"] = cust["c_key"] * 20
    orders["o_cust"] = orders["o_cust"] * 20
    if shuffle_orders:
        perm = np.random.default_rng(7).permutation(len(orders["o_key"]))
        orders = {c: np.ascontiguousarray(v[perm]) for c, v in orders.items()}
    if shuffle_lines:
        perm = np.random.default_rng(8).permutation(len(lines["l_order"]))
        lines = {c: np.ascontiguousarray(v[perm]) for c, v in lines.items()}
    tables = (cust, orders, lines)
    n_c, n_o, n_l = len(cust["c_key"]), len(orders["o_key"]), len(lines["l_order"])

    def build(p):
        d = datagen.star_desc(p, n_c, n_o, n_l, k=50)
        d.d_key_min, d.d_key_max = 20, 20 * n_c
        return d
    rows = _star_parity(ops, tables, build)
    assert len(rows) == 50


@pytest.mark.parametrize("variant", ["float_pred", "no_pred", "negative_keys"])
def test_star_parity_parent_pass_variants(ops, variant):
    """The persistent parent pass's other instantiations (1.5M-row parents):
    a FLOAT parent predicate, no parent predicate, and keys below zero (the
    64-bit key-offset path instead of the 32-bit one)."""
    cust, orders, lines = datagen.star_tables(scale=0.1, workload=33)
    shift = 0
    if variant == "float_pred":
        orders["o_df"] = orders["o_date"].astype(np.float32) + np.float32(0.25)
    if variant == "negative_keys":
        shift = -1_000_000_000
        cust["c_key"] = cust["c_key"] + shift
        orders["o_cust"] = orders["o_cust"] + shift
        orders["o_key"] = orders["o_key"] + shift
        lines["l_order"] = lines["l_order"] + shift
    tables = (cust, orders, lines)
    n_c, n_o, n_l = len(cust["c_key"]), len(orders["o_key"]), len(lines["l_order"])

    def build(p):
        d = datagen.star_desc(p, n_c, n_o, n_l, k=20)
        if variant == "float_pred":
            d.p_preds[0] = abi.make_pred(p("orders", "o_df"), abi.FLOAT, abi.CMP_LT,
                                         float(datagen.STAR_CUTOFF))
        elif variant == "no_pred":
            d.p_n_preds = 0
        else:
            d.d_key_min, d.d_key_max = 1 + shift, n_c + shift
            d.p_key_min, d.p_key_max = 1 + shift, n_o + shift
        return d
    rows = _star_parity(ops, tables, build)
    assert len(rows) == 20


def test_star_no_qualifying_rows(ops):
    tables = datagen.star_tables(scale=0.001, workload=31)
    n_c, n_o, n_l = (len(tables[0]["c_key"]), len(tables[1]["o_key"]),
                     len(tables[2]["l_order"]))
    d, _keep = device_desc(ops, tables, lambda p: datagen.star_desc(p, n_c, n_o, n_l,
                                                                      segment=9))
    assert ops.star_topk(d) == []


def test_star_duplicate_parent_key_falls_back(ops):
    cust = {"c_key": np.array([1], np.int32), "segment": np.array([1], np.uint8)}
    orders = {"o_key": np.array([10, 10], np.int32), "o_cust": np.array([1, 1], np.int32),
              "o_date": np.array([1, 1], np.int32), "o_prio": np.zeros(2, np.int32)}
    lines = {"l_order": np.array([10], np.int32), "price": np.array([1], np.int64),
             "discount": np.zeros(1, np.int64), "l_ship": np.array([50], np.int32)}
    tables = (cust, orders, lines)

    def build(p):
        d = datagen.star_desc(p, 1, 2, 1, segment=1, cutoff=20)
        d.d_key_min = d.d_key_max = 1
        d.p_key_min = d.p_key_max = 10
        return d
    d, _keep = device_desc(ops, tables, build)
    with pytest.raises(abi.Unsupported):
        ops.star_topk(d)


def _one_order_tables(n_lines, price):
    cust = {"c_key": np.array([1], np.int32), "segment": np.array([1], np.uint8)}
    orders = {"o_key": np.array([10], np.int32), "o_cust": np.array([1], np.int32),
              "o_date": np.array([1], np.int32), "o_prio": np.zeros(1, np.int32)}
    lines = {"l_order": np.full(n_lines, 10, np.int32),
             "price": np.full(n_lines, price, np.int64),
             "discount": np.zeros(n_lines, np.int64),
             "l_ship": np.full(n_lines, 50, np.int32)}
    return cust, orders, lines


def _one_order_desc(n_lines):
    def build(p):
        d = datagen.star_desc(p, 1, 1, n_lines, segment=1, cutoff=20)
        d.d_key_min = d.d_key_max = 1
        d.p_key_min = d.p_key_max = 10
        return d
    return build


def test_star_large_sums_exact(ops):
    """Two lines of 100 * 2^55 each: the sum (7.2e18) still fits int64."""
    tables = _one_order_tables(2, 2 ** 55)
    d, _keep = device_desc(ops, tables, _one_order_desc(2))
    rows = ops.star_topk(d)
    assert [(r.p_row, r.count, r.sum_i64) for r in rows] == [(0, 2, 2 * 100 * 2 ** 55)]


def test_star_sum_overflow_bound_falls_back(ops):
    """Three such lines exceed int64: the bound count x max|value| flags it."""
    tables = _one_order_tables(3, 2 ** 55)
    d, _keep = device_desc(ops, tables, _one_order_desc(3))
    with pytest.raises(abi.Unsupported):
        ops.star_topk(d)


def test_star_expression_overflow_is_an_error(ops):
    """price * (100 - discount) itself overflows int64: an overflow error,
    as DuckDB raises for the same expression."""
    tables = _one_order_tables(1, 2 ** 60)
    d, _keep = device_desc(ops, tables, _one_order_desc(1))
    with pytest.raises(abi.MettleError) as ei:
        ops.star_topk(d)
    assert ei.value.code == abi.MTL_ERR_OVERFLOW


def _order_key(col: np.ndarray, t: int) -> np.ndarray:
    """Order-preserving uint64 encoding (DuckDB order: NaN largest, -0 == 0)."""
    if t in (abi.FLOAT, abi.DOUBLE):
        f = col.astype(np.float64)
        f = np.where(f == 0, 0.0, f)
        b = f.view(np.uint64)
        neg = (b >> np.uint64(63)) == 1
        e = np.where(neg, ~b, b | np.uint64(1 << 63))
        return np.where(np.isnan(f), np.uint64(0xFFFFFFFFFFFFFFFF), e)
    return col.astype(np.int64).view(np.uint64) ^ np.uint64(1 << 63)


@pytest.mark.parametrize("t,dtype", [(abi.INTEGER, np.int32), (abi.BIGINT, np.int64),
                                     (abi.DATE, np.int32), (abi.TIMESTAMP, np.int64),
                                     (abi.FLOAT, np.float32), (abi.DOUBLE, np.float64),
                                     (abi.DICT8, np.uint8)])
@pytest.mark.parametrize("desc", [False, True])
def test_sort_parity(ops, t, dtype, desc):
    rng = np.random.default_rng(300 + t + 10 * desc)
    n = 200_003
    if np.issubdtype(dtype, np.floating):
        col = rng.standard_normal(n).astype(dtype)
        col[rng.integers(0, n, 500)] = np.nan
        col[rng.integers(0, n, 500)] = -0.0
        col[rng.integers(0, n, 500)] = 0.0
    else:
        info = np.iinfo(dtype)
        col = rng.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
        col[rng.integers(0, n, 20000)] = col[0]  # many ties
    dcol = ops.to_device(col)
    # sort a selection (every third row) and the whole column
    sel = np.arange(0, n, 3, dtype=np.uint32)
    dsel = ops.to_device(sel.astype(np.int32))
    cnt = torch.tensor([len(sel)], dtype=torch.int32, device="cuda")
    for s, ds, dc in ((sel, dsel, cnt), (np.arange(n, dtype=np.uint32), None, None)):
        k = _order_key(col[s], t)
        if desc:
            k = ~k
        exp = s[np.argsort(k, kind="stable")]
        got, m = ops.sort(dcol, t, n, sel=ds, count=dc, desc=desc)
        assert m == len(s)
        np.testing.assert_array_equal(ops.to_host(got, np.uint32), exp)


@pytest.mark.parametrize("case", ["constant", "high_digits", "narrow", "ties", "multi_tile",
                                  "one_row", "tail"])
def test_sort_radix_pass_shapes(ops, case):
    """The radix sort skips digit positions that are constant over the keys
    and runs the others: every key equal (no pass), keys differing only in
    their high bytes, a narrow BIGINT range (one or two passes of eight),
    many ties; 9M rows (more than one 4096-row tile per block of the 2048
    block ranges); one row; a partial last tile.  Each against numpy's
    stable argsort."""
    rng = np.random.default_rng(["constant", "high_digits", "narrow", "ties", "multi_tile",
                                 "one_row", "tail"].index(case) + 400)
    t, dtype, n = abi.BIGINT, np.int64, 300_001
    if case == "constant":
        col = np.full(n, -17, dtype=np.int64)
    elif case == "high_digits":
        col = (rng.integers(0, 256, n).astype(np.int64) << 56) - (1 << 62)
    elif case == "narrow":
        col = rng.integers(1_000_000, 1_000_300, n).astype(np.int64)
    elif case == "ties":
        col = rng.integers(0, 3, n).astype(np.int64) * (1 << 40)
    elif case == "multi_tile":
        t, dtype, n = abi.INTEGER, np.int32, 9_000_011
        col = rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
    elif case == "one_row":
        col, n = np.array([5], dtype=np.int64), 1
    else:
        t, dtype, n = abi.DATE, np.int32, 4096 * 3 + 17
        col = rng.integers(10957, 22281, n).astype(np.int32)
    dcol = ops.to_device(col)
    for desc in (False, True):
        k = _order_key(col, t)
        if desc:
            k = ~k
        exp = np.argsort(k, kind="stable").astype(np.uint32)
        got, m = ops.sort(dcol, t, n, desc=desc)
        assert m == n
        np.testing.assert_array_equal(ops.to_host(got, np.uint32), exp)


def test_sort_empty_selection(ops):
    col = ops.to_device(np.arange(10, dtype=np.int32))
    sel = ops.to_device(np.zeros(4, dtype=np.int32))
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    got, m = ops.sort(col, abi.INTEGER, 10, sel=sel, count=cnt)
    assert m == 0
