"""Tuned launch parameters: the table (CPU: get/set/save/load without a
device) and, on the GPU, results under every candidate value against the
oracle, and the autotuner itself (it checks every candidate against the
default's result)."""
import ctypes as C

import numpy as np
import pytest

from mettle_amd import abi
from oracle import oracle as orc

from .test_group import compare
from .test_serve import host_fingerprint

F = abi

CANDIDATES = {"small_block": [128, 256, 512, 1024], "small_steps": [1, 2, 4],
              "join_load_pct": [25, 35, 50, 70], "group_load_pct": [25, 35, 50, 70],
              "part_bits": [5, 6, 7, 8, 9, 10]}


def _get(name):
    v = C.c_int32()
    abi.check(abi.lib().mtl_tune_get(name.encode(), C.byref(v)))
    return v.value


@pytest.fixture
def fresh_table():
    abi.check(abi.lib().mtl_tune_reset())
    yield
    abi.check(abi.lib().mtl_tune_reset())


def test_tune_table_defaults_and_candidates(fresh_table):
    L = abi.lib()
    for name, cands in CANDIDATES.items():
        d = _get(name)
        assert d in cands
        for v in cands:
            abi.check(L.mtl_tune_set(name.encode(), v))
            assert _get(name) == v
        # a value outside the candidates is refused and leaves the table as it was
        assert L.mtl_tune_set(name.encode(), 3) == abi.MTL_ERR_INVALID_ARGUMENT
        assert _get(name) == cands[-1]
    assert L.mtl_tune_set(b"no_such_knob", 1) == abi.MTL_ERR_INVALID_ARGUMENT
    buf = C.create_string_buffer(512)
    abi.check(L.mtl_tune_describe(buf, len(buf)))
    desc = dict(kv.split("=") for kv in buf.value.decode().split())
    assert {k: int(v) for k, v in desc.items()} == {k: c[-1] for k, c in CANDIDATES.items()}


def test_tune_save_load_round_trip_and_refusal(tmp_path, fresh_table):
    L = abi.lib()
    abi.check(L.mtl_tune_set(b"join_load_pct", 35))
    abi.check(L.mtl_tune_set(b"part_bits", 9))
    path = str(tmp_path / "tuned.txt").encode()
    abi.check(L.mtl_tune_save(path))
    abi.check(L.mtl_tune_reset())
    assert _get("join_load_pct") == 50
    abi.check(L.mtl_tune_load(path))
    assert (_get("join_load_pct"), _get("part_bits")) == (35, 9)
    # comments and blank lines are skipped
    good = tmp_path / "good.txt"
    good.write_text("# tuned\n\n  small_steps=2\nsmall_block=512\n")
    abi.check(L.mtl_tune_load(str(good).encode()))
    assert (_get("small_steps"), _get("small_block")) == (2, 512)
    # a bad line anywhere refuses the whole file: nothing changes
    for text in ("part_bits=6\nsmall_steps=3\n", "part_bits=6\nbogus=1\n", "part_bits=6\nx\n",
                 "part_bits=six\n"):
        bad = tmp_path / "bad.txt"
        bad.write_text(text)
        assert L.mtl_tune_load(str(bad).encode()) == abi.MTL_ERR_INVALID_ARGUMENT
        assert _get("part_bits") == 9 and _get("small_steps") == 2
    assert L.mtl_tune_load(str(tmp_path / "missing.txt").encode()) == abi.MTL_ERR_INVALID_ARGUMENT


def test_join_table_bytes_follow_the_tuned_fill(fresh_table):
    """The hash join table is sized at the tuned fill: the smallest power of
    two of slots (at least 16) with rows <= fill% of it."""
    L = abi.lib()

    def slots_for(rows, pct):
        c = 16
        while c * pct < rows * 100:
            c *= 2
        return c

    def table_bytes(cap):
        return (((cap + 1) * 8 + 255) & ~255) + (cap + 1) * 4

    for pct in CANDIDATES["join_load_pct"]:
        abi.check(L.mtl_tune_set(b"join_load_pct", pct))
        for rows in (0, 1, 15, 1000, 1 << 20, 3_000_001):
            assert L.mtl_join_table_bytes(abi.JOIN_HASH, rows, 0) == \
                table_bytes(slots_for(rows, pct))


# ---------------------------------------------------------------------------
# GPU: parity under every candidate value
# ---------------------------------------------------------------------------
@pytest.mark.gpu
@pytest.mark.parametrize("pct", CANDIDATES["join_load_pct"])
def test_gpu_hash_join_parity_at_each_fill(pct, fresh_table):
    from mettle_amd import ops
    ops.tune_set("join_load_pct", pct)
    rng = np.random.default_rng(pct)
    # 64-bit keys in a 2^28 span (the oracle's dense join needs a span
    # below 2^31)
    base = -(1 << 40)
    bk = (base + np.unique(rng.integers(0, 1 << 28, 330_000))[:300_000]).astype(np.int64)
    rng.shuffle(bk)
    pk = np.concatenate([bk[rng.integers(0, len(bk), 500_000)],
                         base + rng.integers(-(1 << 27), 1 << 29, 500_000)]).astype(np.int64)
    rng.shuffle(pk)
    ref_p, ref_b = orc.join(bk, abi.BIGINT, pk, abi.BIGINT)
    gp, gb = ops.join(abi.JOIN_HASH, ops.to_device(bk), abi.BIGINT, ops.to_device(pk),
                      abi.BIGINT)
    np.testing.assert_array_equal(ops.to_host(gp, np.uint32), ref_p)
    np.testing.assert_array_equal(ops.to_host(gb, np.uint32), ref_b)


@pytest.mark.gpu
def test_gpu_join_table_built_and_probed_under_different_fills(fresh_table):
    """The probe derives the slots from the table's allocation, so a table
    sized and built at one fill is probed correctly after the fill changes."""
    import torch
    from mettle_amd import ops
    L = abi.lib()
    rng = np.random.default_rng(5)
    bk = np.unique(rng.integers(0, 1 << 28, 110_000))[:100_000].astype(np.int64)
    rng.shuffle(bk)
    pk = bk[rng.integers(0, len(bk), 200_000)]
    ops.tune_set("join_load_pct", 25)
    nbytes = L.mtl_join_table_bytes(abi.JOIN_HASH, len(bk), 0)
    table = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    dup = torch.zeros(4, dtype=torch.int32, device="cuda")
    dbk, dpk = ops.to_device(bk), ops.to_device(pk)
    abi.check(L.mtl_join_build(abi.JOIN_HASH, dbk.data_ptr(), abi.BIGINT, None, None, len(bk), 0,
                               0, table.data_ptr(), nbytes, dup.data_ptr(), None))
    ops.tune_set("join_load_pct", 70)
    op = torch.empty(len(pk), dtype=torch.int32, device="cuda")
    ob = torch.empty(len(pk), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(4, dtype=torch.int32, device="cuda")
    abi.check(L.mtl_join_probe(abi.JOIN_HASH, table.data_ptr(), nbytes, 0, len(bk),
                               dpk.data_ptr(), abi.BIGINT, None, None, len(pk), op.data_ptr(),
                               ob.data_ptr(), cnt.data_ptr(), None))
    torch.cuda.synchronize()
    n = int(cnt[0])
    ref_p, ref_b = orc.join(bk, abi.BIGINT, pk, abi.BIGINT)
    assert n == len(ref_p)
    np.testing.assert_array_equal(op[:n].cpu().numpy().astype(np.uint32), ref_p)
    np.testing.assert_array_equal(ob[:n].cpu().numpy().astype(np.uint32), ref_b)


def _group_case(n, keys, seed):
    rng = np.random.default_rng(seed)
    vi = rng.integers(-10**9, 10**9 + 1, n).astype(np.int64)
    aggs = [(F.AGG_SUM, vi, F.BIGINT), (F.AGG_COUNT, None, 0), (F.AGG_MIN, vi, F.BIGINT),
            (F.AGG_MAX, vi, F.BIGINT)]
    return keys, vi, aggs


def _group_check(keys, vi, aggs, kmin, kmax):
    from mettle_amd import ops
    rk, ro = orc.group_agg([(keys, F.BIGINT)], aggs)
    dk, dv = ops.to_device(keys), ops.to_device(vi)
    g = ops.group_agg([(dk, F.BIGINT, kmin, kmax)],
                      [(f, None if c is None else dv, t) for f, c, t in aggs], len(keys))
    gk, go = g.fetch()
    # a hash table's groups come in table order: sort by key like the oracle
    o = np.argsort(gk[:, 0], kind="stable")
    compare(gk[o], [x[o] for x in go], rk, ro, aggs)


@pytest.mark.gpu
@pytest.mark.parametrize("pct", CANDIDATES["group_load_pct"])
def test_gpu_group_hash_table_parity_at_each_fill(pct, fresh_table):
    from mettle_amd import ops
    ops.tune_set("group_load_pct", pct)
    rng = np.random.default_rng(100 + pct)
    uniq = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, 200_000,
                        dtype=np.int64)
    keys = uniq[rng.integers(0, len(uniq), 1_500_000)]
    keys, vi, aggs = _group_case(len(keys), keys, pct)
    _group_check(keys, vi, aggs, np.iinfo(np.int64).min, np.iinfo(np.int64).max)


@pytest.mark.gpu
@pytest.mark.parametrize("bits", CANDIDATES["part_bits"])
def test_gpu_group_spill_partition_parity_at_each_fan_out(bits, fresh_table):
    """A skewed key over a 3M-key dense domain (2.5M rows: the cache misses
    are spilled and partitioned in levels of at most `bits` bits)."""
    from mettle_amd import ops
    ops.tune_set("part_bits", bits)
    rng = np.random.default_rng(200 + bits)
    n = 2_500_000
    keys = (rng.zipf(1.2, n) % 3_000_000).astype(np.int64)
    keys, vi, aggs = _group_case(n, keys, bits)
    _group_check(keys, vi, aggs, 0, 2_999_999)


@pytest.mark.gpu
@pytest.mark.parametrize("block,steps", [(b, s) for b in CANDIDATES["small_block"]
                                         for s in CANDIDATES["small_steps"]])
def test_gpu_small_scans_parity_at_each_block_shape(block, steps, fresh_table):
    """Small scans through a server whose graph was captured with the given
    block size and steps per wave, against the oracle."""
    from mettle_amd import ops
    ops.tune_set("small_block", block)
    ops.tune_set("small_steps", steps)
    rng = np.random.default_rng(block + steps)
    n_max = 1 << 16
    v = rng.integers(0, 2**20, n_max).astype(np.int32)
    b = rng.integers(-2**40, 2**40, n_max).astype(np.int64)
    dv, db = ops.to_device(v), ops.to_device(b)
    work = []
    for j in range(40):
        n = int(10 ** rng.uniform(0, np.log10(n_max)))
        lit = int(v[int(rng.integers(0, n_max))])
        preds = [(dv, abi.INTEGER, abi.CMP_GT, lit)]
        hp = [(v, abi.INTEGER, abi.CMP_GT, lit)]
        if j % 2:
            preds.append((db, abi.BIGINT, abi.CMP_LT, 0))
            hp.append((b, abi.BIGINT, abi.CMP_LT, 0))
        q = ops.scan_query(preds, n, (db, abi.BIGINT, abi.PROJ_ADD, 7))
        sel = orc.filter(hp, n)
        work.append((q, (len(sel), host_fingerprint(sel, (b[sel] + 7).view(np.uint64)))))
    with ops.Server(1, budget=64 << 30, max_batch=64) as srv:
        t0 = abi.lib().mtl_now_ns() + 40_000_000
        ts = [srv.submit(q, t0) for q, _ in work]
        res = [srv.wait(t) for t in ts]
    assert max(r.batch for r in res) > 1
    for (q, exp), r in zip(work, res):
        assert (r.count, r.fingerprint) == exp


@pytest.mark.gpu
def test_gpu_autotune_runs_and_keeps_candidates(fresh_table):
    from mettle_amd import ops
    rep = ops.autotune(ops.TUNE_ALL)
    chosen = [ln for ln in rep.splitlines() if ln.endswith("(chosen)")]
    assert len(chosen) == 4  # small_scan, hash_join, group_hash, group_spill
    t = ops.tune_describe()
    for name, cands in CANDIDATES.items():
        assert t[name] in cands
