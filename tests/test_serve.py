"""Serving loop (mtl_server): every query's result is the same alone and in
a concurrent mix, scan results match the oracle, admission honours the
scratch budget, and scheduled arrivals are respected."""
import random

import numpy as np
import pytest

from mettle_amd import abi, datagen
from oracle import oracle as orc

pytestmark = pytest.mark.gpu



def host_fingerprint(rows: np.ndarray, vals_bits: np.ndarray) -> int:
    """The device fingerprint restated: sum over positions i of a 32-bit
    multiply-xorshift of (i, row_i, value bits_i), in uint64."""
    i = np.arange(len(rows), dtype=np.uint64).astype(np.uint32)
    v = vals_bits.astype(np.uint64)
    with np.errstate(over="ignore"):
        x = i * np.uint32(0x9E3779B1) + rows.astype(np.uint32)
        x ^= (v & np.uint64(0xFFFFFFFF)).astype(np.uint32) * np.uint32(0x85EBCA77)
        x ^= (v >> np.uint64(32)).astype(np.uint32) * np.uint32(0xC2B2AE3D)
        x ^= x >> np.uint32(15)
        x *= np.uint32(0x2C1B3C6D)
        x ^= x >> np.uint32(13)
        return int(x.astype(np.uint64).sum(dtype=np.uint64))


@pytest.fixture(scope="module")
def ops():
    from mettle_amd import ops as _ops
    return _ops


def _scan_workload(ops, rng, n):
    # device columns hold at least one row (n may be 0)
    v = rng.integers(0, 2**20, max(n, 1)).astype(np.int32)
    b = rng.integers(-2**40, 2**40, max(n, 1)).astype(np.int64)
    dv, db = ops.to_device(v), ops.to_device(b)
    v, b = v[:n], b[:n]
    cut = int(rng.integers(0, 2**20))
    q1 = ops.scan_query([(dv, abi.INTEGER, abi.CMP_GT, cut)], n, (dv, abi.INTEGER, abi.PROJ_MUL, 10))
    sel = orc.filter([(v, abi.INTEGER, abi.CMP_GT, cut)], n)
    exp1 = (len(sel), host_fingerprint(sel, (v[sel].astype(np.int64) * 10).astype(np.uint32)))
    q2 = ops.scan_query([(db, abi.BIGINT, abi.CMP_LT, 0), (dv, abi.INTEGER, abi.CMP_LE, cut)], n,
                        (db, abi.BIGINT, abi.PROJ_ADD, 7))
    sel2 = orc.filter([(b, abi.BIGINT, abi.CMP_LT, 0), (v, abi.INTEGER, abi.CMP_LE, cut)], n)
    exp2 = (len(sel2), host_fingerprint(sel2, (b[sel2] + 7).view(np.uint64)))
    # no projection: the value of each position is its row id
    q3 = ops.scan_query([(db, abi.BIGINT, abi.CMP_GE, -2**39)], n)
    sel3 = orc.filter([(b, abi.BIGINT, abi.CMP_GE, -2**39)], n)
    exp3 = (len(sel3), host_fingerprint(sel3, sel3))
    return [(q1, exp1), (q2, exp2), (q3, exp3)], (dv, db)


def test_scan_results_match_oracle_alone_and_in_mix(ops):
    rng = np.random.default_rng(5)
    work, keep = [], []
    for n in (0, 1, 4095, 4097, 100_003, 1_000_000):
        w, k = _scan_workload(ops, rng, n)
        work += w
        keep.append(k)
    # alone: one slot, one query at a time
    with ops.Server(1) as srv:
        alone = []
        for q, exp in work:
            r = srv.wait(srv.submit(q))
            assert (r.count, r.fingerprint) == exp
            alone.append((r.count, r.fingerprint))
    # mixed: eight slots, shuffled, each query submitted three times
    order = [i for i in range(len(work)) for _ in range(3)]
    random.Random(1).shuffle(order)
    with ops.Server(8) as srv:
        tickets = [(i, srv.submit(work[i][0])) for i in order]
        for i, t in tickets:
            r = srv.wait(t)
            assert (r.count, r.fingerprint) == alone[i]


def test_star_results_same_alone_and_in_mix(ops):
    sizes = [0.001, 0.003, 0.01]
    descs, keep, expect = [], [], []
    for j, sc in enumerate(sizes):
        tables = datagen.star_tables(scale=sc, workload=40 + j)
        n_c, n_o, n_l = (len(tables[0]["c_key"]), len(tables[1]["o_key"]),
                         len(tables[2]["l_order"]))
        dev = {name: {c: ops.to_device(v) for c, v in t.items()}
               for name, t in zip(["customers", "orders", "lines"], tables)}
        d = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, n_o, n_l)
        by = dict(zip(["customers", "orders", "lines"], tables))
        st, ref = orc.star_topk(datagen.star_desc(lambda t, c: by[t][c].ctypes.data, n_c, n_o, n_l))
        assert st == abi.MTL_OK
        descs.append(d)
        keep.append((dev, tables))
        expect.append([(r.p_row, r.count, r.sum_i64) for r in ref])
    scan_work, scan_keep = _scan_workload(ops, np.random.default_rng(9), 300_000)
    with ops.Server(6) as srv:
        pending = []
        for rep in range(4):
            for j, d in enumerate(descs):
                q, out = ops.star_query(d)
                pending.append(("star", j, q, out, srv.submit(q)))
            for q, exp in scan_work:
                pending.append(("scan", exp, q, None, srv.submit(q)))
        for kind, info, q, out, t in pending:
            r = srv.wait(t)
            if kind == "star":
                got = [(out[i].p_row, out[i].count, out[i].sum_i64) for i in range(r.n_rows)]
                assert got == expect[info]
            else:
                assert (r.count, r.fingerprint) == info


def test_admission_budget(ops):
    rng = np.random.default_rng(11)
    work, keep = _scan_workload(ops, rng, 200_000)
    need = abi.lib().mtl_query_scratch_bytes(work[0][0])
    assert need > 200_000 * 4
    # a budget smaller than the query: refused with MTL_ERR_BUDGET
    with ops.Server(4, budget=need // 2) as srv:
        t = srv.submit(work[0][0])
        with pytest.raises(abi.MettleError) as ei:
            srv.wait(t)
        assert ei.value.code == abi.MTL_ERR_BUDGET
    # room for one query at a time: all complete, with correct results
    big = max(abi.lib().mtl_query_scratch_bytes(q) for q, _ in work)
    with ops.Server(4, budget=big) as srv:
        ts = [(exp, srv.submit(q)) for q, exp in work for _ in range(4)]
        res = [(exp, srv.wait(t)) for exp, t in ts]
        for exp, r in res:
            assert (r.count, r.fingerprint) == exp
        # admitted one at a time: no two runs overlap
        spans = sorted((r.t_start_ns, r.t_done_ns) for _, r in res)
        for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
            assert s1 >= e0 - 1_000_000  # 1 ms slack for the host stamps


def test_scheduled_arrivals(ops):
    rng = np.random.default_rng(13)
    work, keep = _scan_workload(ops, rng, 50_000)
    lib = abi.lib()
    with ops.Server(2) as srv:
        t0 = lib.mtl_now_ns() + 20_000_000
        arrivals = [t0 + k * 2_000_000 for k in range(6)]
        ts = [srv.submit(work[k % 2][0], a) for k, a in enumerate(arrivals)]
        for a, t in zip(arrivals, ts):
            r = srv.wait(t)
            assert r.t_submit_ns == a
            assert r.t_start_ns >= a
            assert r.t_done_ns >= r.t_start_ns
