"""Continuous batching in the serving loop: queries that arrive together run
as one batch whose scans and star joins share their table passes.  Every
query's result in a batch equals the oracle's and its result alone (serial
dispatch); SQL and group-by queries enter the same loop; FLOAT sums agree
alone and in a mix within the stated tolerance; the processing budget of
gpu_buffer_init refuses what does not fit and releases what is freed."""
import ctypes as C

import numpy as np
import pytest

from mettle_amd import abi, datagen
from mettle_amd import sql as S
from oracle import oracle as orc

from .test_serve import host_fingerprint

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

# FLOAT/DOUBLE sums: |a - b| <= FLOAT_REL * sum(|v|) over the group's rows
FLOAT_REL = 1e-6
BUDGET = 64 << 30  # admission budget of the servers here (device scratch)


@pytest.fixture(scope="module")
def ops():
    from mettle_amd import ops as _ops
    return _ops


def _arrive_together(srv, queries, lead_ns=40_000_000):
    """Submit every query with one arrival time a little ahead, so the
    workers see them all arrived at once and form batches."""
    t0 = abi.lib().mtl_now_ns() + lead_ns
    return [srv.submit(q, t0) for q in queries]


def _bits(vals: np.ndarray) -> np.ndarray:
    """The fingerprint's value bits: 4-byte values zero-extended."""
    if vals.dtype.itemsize == 4:
        return vals.view(np.uint32)
    return vals.view(np.uint64)


def _scan_set(ops, rng, n_max):
    v = rng.integers(0, 2**20, n_max).astype(np.int32)
    b = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n_max, dtype=np.int64,
                     endpoint=True)
    f = rng.standard_normal(n_max).astype(np.float32)
    dev = {"v": ops.to_device(v), "b": ops.to_device(b), "f": ops.to_device(f)}
    host = {"v": v, "b": b, "f": f}
    types = {"v": abi.INTEGER, "b": abi.BIGINT, "f": abi.FLOAT}
    work = []
    for j in range(40):
        col = ["v", "v", "b", "f"][j % 4]
        t = types[col]
        n = int(10 ** rng.uniform(0, np.log10(n_max))) if j % 7 else n_max
        op = int(rng.integers(0, 5))
        x = host[col]
        lit = float(x[int(rng.integers(0, n_max))]) if t == abi.FLOAT else int(
            x[int(rng.integers(0, n_max))])
        if col == "v":
            proj = [(abi.PROJ_MUL, 10), (abi.PROJ_ADD, 3), None][j % 3]
        elif col == "b":
            proj = [(abi.PROJ_SUB, 1), None][j % 2]
        else:
            proj = (abi.PROJ_MUL, 2.5)
        q = ops.scan_query([(dev[col], t, op, lit)], n,
                           None if proj is None else (dev[col], t, proj[0], proj[1]))
        sel = orc.filter([(x, t, op, lit)], n)
        if proj is None:
            exp = (len(sel), host_fingerprint(sel, sel), False)
        else:
            vals, ovf = orc.project(x, t, sel, proj[0], proj[1])
            exp = (len(sel), None if ovf else host_fingerprint(sel, _bits(vals)), ovf)
        work.append((q, exp))
    # not shareable: two predicates (runs alone inside the batch)
    q = ops.scan_query([(dev["b"], abi.BIGINT, abi.CMP_LT, 0),
                        (dev["v"], abi.INTEGER, abi.CMP_GE, 1000)], n_max,
                       (dev["b"], abi.BIGINT, abi.PROJ_ADD, 7))
    sel = orc.filter([(b, abi.BIGINT, abi.CMP_LT, 0), (v, abi.INTEGER, abi.CMP_GE, 1000)], n_max)
    work.append((q, (len(sel), host_fingerprint(sel, (b[sel] + 7).view(np.uint64)), False)))
    return work, dev


def _check_scan(r, exp):
    n, fp, ovf = exp
    assert r.count == n
    if fp is not None:
        assert r.fingerprint == fp


def _run(ops, work, n_slots, max_batch, together=True):
    """Results (or the exception) of every query, in submission order."""
    out = []
    with ops.Server(n_slots, budget=BUDGET, max_batch=max_batch) as srv:
        qs = [q for q, _ in work]
        ts = _arrive_together(srv, qs) if together else [srv.submit(q) for q in qs]
        for t in ts:
            try:
                out.append(srv.wait(t))
            except abi.MettleError as e:
                out.append(e)
    return out


def test_shared_scan_batch_matches_oracle_and_alone(ops):
    rng = np.random.default_rng(21)
    work, keep = _scan_set(ops, rng, 3_000_000)
    serial = _run(ops, work, 1, 1, together=False)
    batched = _run(ops, work, 2, 32)
    assert max(r.batch for r in batched if isinstance(r, abi.QueryResult)) > 1
    for (q, exp), a, b in zip(work, serial, batched):
        if exp[2]:  # the oracle overflowed: both raise MTL_ERR_OVERFLOW
            assert isinstance(a, abi.MettleError) and a.code == abi.MTL_ERR_OVERFLOW
            assert isinstance(b, abi.MettleError) and b.code == abi.MTL_ERR_OVERFLOW
            continue
        _check_scan(a, exp)
        _check_scan(b, exp)
        assert (a.count, a.fingerprint) == (b.count, b.fingerprint)


def test_small_scans_in_the_persistent_graph_launch(ops):
    """Small scans (at most 2^16 rows) of every column type with one to
    three predicates and projections run in the batch's persistent launch,
    replayed from the slot's HIP graph: three batches in a row on one
    server, with eager device work and host synchronisation between them.
    Every result equals the oracle's (count and fingerprint, or the
    overflow error)."""
    import torch
    rng = np.random.default_rng(23)
    n_max = 1 << 16
    cols = {
        "i": (rng.integers(-2**31, 2**31, n_max).astype(np.int32), abi.INTEGER),
        "b": (rng.integers(-2**62, 2**62, n_max).astype(np.int64), abi.BIGINT),
        "f": (rng.standard_normal(n_max).astype(np.float32), abi.FLOAT),
        "g": (rng.standard_normal(n_max), abi.DOUBLE),
        "d": (rng.integers(10957, 22281, n_max).astype(np.int32), abi.DATE),
        "t": (rng.integers(946684800 * 10**6, 1924991999 * 10**6, n_max).astype(np.int64),
              abi.TIMESTAMP),
        "c": (rng.integers(0, 5, n_max).astype(np.uint8), abi.DICT8),
    }
    dev = {k: ops.to_device(x) for k, (x, _) in cols.items()}
    projs = {"i": [(abi.PROJ_MUL, 3), (abi.PROJ_ADD, 2**30)], "b": [(abi.PROJ_SUB, 5)],
             "f": [(abi.PROJ_MUL, 0.5)], "g": [(abi.PROJ_ADD, 1.25)],
             "d": [(abi.PROJ_ADD, 30)], "t": [(abi.PROJ_SUB, 10**6)], "c": [None]}
    names = list(cols)
    rounds = []
    for _ in range(3):
        work = []
        for j in range(48):
            n = int(10 ** rng.uniform(0, np.log10(n_max))) if j % 5 else n_max
            preds = []
            for k in range(1 + j % 3):
                c = names[int(rng.integers(0, len(names)))]
                x, t = cols[c]
                lit = x[int(rng.integers(0, n_max))]
                lit = float(lit) if t in (abi.FLOAT, abi.DOUBLE) else int(lit)
                preds.append((c, t, int(rng.integers(0, 5)), lit))
            pc = names[j % len(names)]
            pr = projs[pc][j % len(projs[pc])]
            x, t = cols[pc]
            q = ops.scan_query([(dev[c], t, op, lit) for c, t, op, lit in preds], n,
                               None if pr is None else (dev[pc], t, pr[0], pr[1]))
            sel = orc.filter([(cols[c][0], t, op, lit) for c, t, op, lit in preds], n)
            if pr is None:
                exp = (len(sel), host_fingerprint(sel, sel), False)
            else:
                vals, ovf = orc.project(x, t, sel, pr[0], pr[1])
                exp = (len(sel), None if ovf else host_fingerprint(sel, _bits(vals)), ovf)
            work.append((q, exp))
        rounds.append(work)
    with ops.Server(1, budget=BUDGET, max_batch=64) as srv:
        for work in rounds:
            ts = _arrive_together(srv, [q for q, _ in work])
            for (q, exp), t in zip(work, ts):
                try:
                    r = srv.wait(t)
                except abi.MettleError as e:
                    assert exp[2] and e.code == abi.MTL_ERR_OVERFLOW
                    continue
                assert not exp[2]
                assert r.batch > 1
                _check_scan(r, exp)
            # eager device work and a synchronisation between the replays
            z = torch.arange(1 << 20, device="cuda").sum()
            torch.cuda.synchronize()
            assert int(z) == (1 << 20) * ((1 << 20) - 1) // 2


def test_shared_scan_batch_at_full_table_size(ops):
    """32 scans over one resident 100M-row INTEGER column in one batch."""
    rng = np.random.default_rng(22)
    n = 100_000_000
    v = rng.integers(0, 2**20, n).astype(np.int32)
    dv = ops.to_device(v)
    work = []
    for j in range(32):
        lit = int(rng.integers(0, 2**20))
        rows = n if j < 4 else int(10 ** rng.uniform(4, 8))
        q = ops.scan_query([(dv, abi.INTEGER, abi.CMP_GT, lit)], rows,
                           (dv, abi.INTEGER, abi.PROJ_MUL, 10))
        sel = orc.filter([(v, abi.INTEGER, abi.CMP_GT, lit)], rows)
        work.append((q, (len(sel), host_fingerprint(sel, (v[sel] * 10).view(np.uint32)), False)))
    res = _run(ops, work, 1, 32)
    assert max(r.batch for r in res) == 32
    for (q, exp), r in zip(work, res):
        _check_scan(r, exp)


def _star_set(ops, scale=0.02, workload=60):
    tables = datagen.star_tables(scale=scale, workload=workload)
    cust, orders, lines = tables
    dev = {name: {c: ops.to_device(x) for c, x in t.items()}
           for name, t in zip(["customers", "orders", "lines"], tables)}
    by = dict(zip(["customers", "orders", "lines"], tables))
    n_c, n_o, n_l = len(cust["c_key"]), len(orders["o_key"]), len(lines["l_order"])
    rng = np.random.default_rng(workload)
    work = []
    for j in range(24):
        nl = n_l if j % 5 == 0 else int(rng.integers(1, n_l + 1))
        no = int(lines["l_order"][nl - 1])  # orders 1..no cover the first nl lines
        seg = int(rng.integers(0, 5))
        cut = datagen.DAYS_1992_01_01 + int(rng.integers(0, datagen.SPAN_7Y))
        k = [10, 100, 1][j % 3]
        dd = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, no, nl, k=k,
                               segment=seg, cutoff=cut)
        hd = datagen.star_desc(lambda t, c: by[t][c].ctypes.data, n_c, no, nl, k=k,
                               segment=seg, cutoff=cut)
        st, ref = orc.star_topk(hd)
        assert st == abi.MTL_OK
        q, out = ops.star_query(dd)
        work.append((q, out, [(r.p_row, r.count, r.sum_i64) for r in ref]))
    return work, dev


def _star_rows(r, out):
    return [(out[i].p_row, out[i].count, out[i].sum_i64) for i in range(r.n_rows)]


def test_shared_star_batch_matches_oracle_and_alone(ops):
    work, keep = _star_set(ops)
    for n_slots, max_batch in ((1, 1), (2, 16), (1, 32)):
        with ops.Server(n_slots, budget=BUDGET, max_batch=max_batch) as srv:
            ts = _arrive_together(srv, [q for q, _, _ in work])
            res = [srv.wait(t) for t in ts]
        if max_batch > 1:
            assert max(r.batch for r in res) > 1
        for (q, out, exp), r in zip(work, res):
            assert _star_rows(r, out) == exp


def test_shared_star_batch_dimension_slots(ops):
    """Sixty-four star queries in one batch (one share group's limit) whose
    dimension predicates take twelve distinct (operator, literal) forms: a
    share group holds at most eight distinct ones, so the batch splits, and
    queries with the same dimension predicate share one dimension bitmap.
    Every result equals the oracle's."""
    tables = datagen.star_tables(scale=0.01, workload=61)
    cust, orders, lines = tables
    dev = {name: {c: ops.to_device(x) for c, x in t.items()}
           for name, t in zip(["customers", "orders", "lines"], tables)}
    by = dict(zip(["customers", "orders", "lines"], tables))
    n_c, n_l = len(cust["c_key"]), len(lines["l_order"])
    rng = np.random.default_rng(61)
    forms = [(op, lit) for op in (abi.CMP_EQ, abi.CMP_LT, abi.CMP_GE) for lit in range(4)]
    work = []
    for j in range(64):
        nl = n_l if j % 7 == 0 else int(rng.integers(1, n_l + 1))
        no = int(lines["l_order"][nl - 1])
        cut = datagen.DAYS_1992_01_01 + int(rng.integers(0, datagen.SPAN_7Y))
        op, lit = forms[j % len(forms)]
        descs = []
        for ptr in (lambda t, c: dev[t][c].data_ptr(), lambda t, c: by[t][c].ctypes.data):
            d = datagen.star_desc(ptr, n_c, no, nl, k=[10, 1, 100][j % 3], cutoff=cut)
            d.d_preds[0] = abi.make_pred(ptr("customers", "segment"), abi.DICT8, op, lit)
            descs.append(d)
        st, ref = orc.star_topk(descs[1])
        assert st == abi.MTL_OK
        q, out = ops.star_query(descs[0])
        work.append((q, out, [(r.p_row, r.count, r.sum_i64) for r in ref]))
    with ops.Server(1, budget=BUDGET, max_batch=64) as srv:
        ts = _arrive_together(srv, [q for q, _, _ in work])
        res = [srv.wait(t) for t in ts]
    assert max(r.batch for r in res) > 8
    for (q, out, exp), r in zip(work, res):
        assert _star_rows(r, out) == exp


def test_shared_star_batch_full_size(ops):
    """Eight star queries over the full 60M-line tables in one batch."""
    work, keep = [], None
    tables = datagen.star_tables(scale=1.0)
    cust, orders, lines = tables
    dev = {name: {c: ops.to_device(x) for c, x in t.items()}
           for name, t in zip(["customers", "orders", "lines"], tables)}
    by = dict(zip(["customers", "orders", "lines"], tables))
    n_c, n_o, n_l = len(cust["c_key"]), len(orders["o_key"]), len(lines["l_order"])
    for j in range(8):
        seg, cut = j % 5, datagen.DAYS_1992_01_01 + 300 * (j + 1)
        dd = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, n_o, n_l, segment=seg,
                               cutoff=cut)
        st, ref = orc.star_topk(datagen.star_desc(lambda t, c: by[t][c].ctypes.data, n_c, n_o,
                                                  n_l, segment=seg, cutoff=cut))
        assert st == abi.MTL_OK
        q, out = ops.star_query(dd)
        work.append((q, out, [(r.p_row, r.count, r.sum_i64) for r in ref]))
    with ops.Server(1, budget=BUDGET, max_batch=8) as srv:
        ts = _arrive_together(srv, [q for q, _, _ in work])
        res = [srv.wait(t) for t in ts]
    assert res[0].batch == 8
    for (q, out, exp), r in zip(work, res):
        assert _star_rows(r, out) == exp


def test_shared_star_batch_mixed_sizes(ops):
    """One batch of sixteen star queries over 7.2M-line tables, half of them
    on more than 4M lines and half on 1K to 4M (log-uniform), with K of 1, 10
    and 100: every result equals the oracle's."""
    tables = datagen.star_tables(scale=0.12, workload=62)
    cust, orders, lines = tables
    dev = {name: {c: ops.to_device(x) for c, x in t.items()}
           for name, t in zip(["customers", "orders", "lines"], tables)}
    by = dict(zip(["customers", "orders", "lines"], tables))
    n_c, n_l = len(cust["c_key"]), len(lines["l_order"])
    assert n_l > (4 << 20) + 100_000
    rng = np.random.default_rng(62)
    work, big = [], 0
    for j in range(16):
        if j % 2:
            nl = int(rng.integers((4 << 20) + 1, n_l + 1))
        else:
            nl = int(10 ** rng.uniform(3, np.log10(4 << 20)))
        big += nl >= (4 << 20)
        no = int(lines["l_order"][nl - 1])
        seg = int(rng.integers(0, 5))
        cut = datagen.DAYS_1992_01_01 + int(rng.integers(0, datagen.SPAN_7Y))
        k = [10, 100, 1][j % 3]
        dd = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, no, nl, k=k,
                               segment=seg, cutoff=cut)
        st, ref = orc.star_topk(datagen.star_desc(lambda t, c: by[t][c].ctypes.data, n_c, no,
                                                  nl, k=k, segment=seg, cutoff=cut))
        assert st == abi.MTL_OK
        q, out = ops.star_query(dd)
        work.append((q, out, [(r.p_row, r.count, r.sum_i64) for r in ref]))
    assert big == 8
    with ops.Server(1, budget=BUDGET, max_batch=16) as srv:
        ts = _arrive_together(srv, [q for q, _, _ in work])
        res = [srv.wait(t) for t in ts]
    assert max(r.batch for r in res) > 8
    for (q, out, exp), r in zip(work, res):
        assert _star_rows(r, out) == exp


def _register_numbers(ops, v):
    dv = ops.to_device(v)
    S.register_table("numbers", len(v), [("value", abi.INTEGER, dv.data_ptr(), 0, 2**20 - 1,
                                          False)])
    return dv


def test_sql_and_group_queries_in_the_serving_loop(ops):
    """The config-1 SQL, a grouped top-K in SQL, group-by descriptors and
    shared scans in one stream of arrivals: each result equals its result
    alone; an unsupported statement completes with MTL_UNSUPPORTED."""
    S.drop()
    try:
        v = datagen.numbers_table()["value"]
        dv = _register_numbers(ops, v)
        rng = np.random.default_rng(31)
        n = 400_000
        g = rng.integers(0, 3000, n).astype(np.int64)
        w = rng.integers(-10**9, 10**9, n).astype(np.int64)
        f = rng.standard_normal(n).astype(np.float32)
        dg, dw, df = ops.to_device(g), ops.to_device(w), ops.to_device(f)
        S.register_table("t", n, [("g", abi.BIGINT, dg.data_ptr(), 0, 2999, False),
                                  ("w", abi.BIGINT, dw.data_ptr(), -10**9, 10**9, False)])
        sqls = ["SELECT value, value * 10 AS scaled FROM numbers WHERE value > 524288 "
                "ORDER BY value",
                "SELECT g, SUM(w) AS s, COUNT(*) AS c FROM t GROUP BY g ORDER BY s DESC LIMIT 20",
                "SELECT value FROM numbers WHERE value < 1000 ORDER BY value DESC LIMIT 5"]
        alone = [S.execute(q) for q in sqls]
        gq = [ops.group_query([(dg, abi.BIGINT, 0, 2999)],
                              [(abi.AGG_SUM, dw, abi.BIGINT), (abi.AGG_COUNT, None, 0)], n,
                              k=50, agg=0, desc=bool(j % 2)) for j in range(4)]
        rk, ro = orc.group_agg([(g, abi.BIGINT)], [(abi.AGG_SUM, w, abi.BIGINT),
                                                    (abi.AGG_COUNT, None, 0)])
        sums = [abi.hugeint_to_int(r) for r in ro[0]]
        scans = [ops.scan_query([(dv, abi.INTEGER, abi.CMP_GT, int(c))], len(v),
                                (dv, abi.INTEGER, abi.PROJ_MUL, 10))
                 for c in rng.integers(0, 2**20, 8)]
        with ops.Server(2, max_batch=16) as srv:
            t0 = abi.lib().mtl_now_ns() + 40_000_000
            st = [srv.submit_sql(q, t0) for q in sqls for _ in range(3)]
            gt = [srv.submit(q, t0) for q, _ in gq]
            ct = [srv.submit(q, t0) for q in scans]
            bad = srv.submit_sql("SELECT value FROM numbers WHERE value NOT IN (1, 2)", t0)
            got_sql = [srv.wait_sql(t) for t in st]
            got_g = [srv.wait(t) for t in gt]
            got_c = [srv.wait(t) for t in ct]
            with pytest.raises(abi.Unsupported):
                srv.wait(bad)
        for i, out in enumerate(got_sql):
            exp = alone[i // 3]
            assert list(out) == list(exp)
            for col in exp:
                assert out[col][:2] == exp[col][:2]
                a, b = out[col][2], exp[col][2]
                assert list(a) == list(b)
        for j, ((q, (kout, outs)), r) in enumerate(zip(gq, got_g)):
            desc = bool(j % 2)
            order = sorted(range(len(rk)), key=lambda i: ((-1 if desc else 1) * sums[i],
                                                          rk[i, 0]))[:50]
            assert r.n_rows == 50 and r.groups == len(rk)
            assert kout[:, 0].tolist() == [int(rk[i, 0]) for i in order]
            assert [abi.hugeint_to_int(x) for x in outs[0]] == [sums[i] for i in order]
        for q, r in zip(scans, got_c):
            lit = q.preds[0].scalar
            sel = np.flatnonzero(v > lit).astype(np.uint32)
            assert (r.count, r.fingerprint) == (len(sel), host_fingerprint(
                sel, (v[sel] * 10).view(np.uint32)))
    finally:
        S.drop()


def test_float_group_sums_alone_and_in_mix(ops):
    """SUM/AVG over a FLOAT column run in double with atomics: alone and in a
    concurrent mix the results agree within FLOAT_REL of sum(|v|), and both
    agree with the oracle's compensated sum within the same bound."""
    rng = np.random.default_rng(41)
    n = 2_000_000
    g = rng.integers(0, 1000, n).astype(np.int64)
    f = (rng.standard_normal(n) * 1e3).astype(np.float32)
    dg, df = ops.to_device(g), ops.to_device(f)
    aggs = [(abi.AGG_SUM, df, abi.FLOAT), (abi.AGG_AVG, df, abi.FLOAT), (abi.AGG_COUNT, None, 0)]

    def make():
        return ops.group_query([(dg, abi.BIGINT, 0, 999)], aggs, n, capacity=1000)
    alone_q, (ak, ao) = make()
    with ops.Server(1) as srv:
        srv.wait(srv.submit(alone_q))
    mix = [make() for _ in range(6)]
    v = rng.integers(0, 2**20, 4_000_000).astype(np.int32)
    dv = ops.to_device(v)
    scans = [ops.scan_query([(dv, abi.INTEGER, abi.CMP_LT, int(c))], len(v))
             for c in rng.integers(0, 2**20, 6)]
    with ops.Server(4, max_batch=8) as srv:
        ts = _arrive_together(srv, [q for q, _ in mix] + scans)
        for t in ts:
            srv.wait(t)
    rk, ro = orc.group_agg([(g, abi.BIGINT)], [(abi.AGG_SUM, f, abi.FLOAT),
                                                (abi.AGG_AVG, f, abi.FLOAT)])
    fa = np.bincount(g, weights=np.abs(f.astype(np.float64)))[rk[:, 0]]
    cnt = np.bincount(g)[rk[:, 0]]
    np.testing.assert_array_equal(ak[:, 0], rk[:, 0])
    assert np.all(np.abs(ao[0] - ro[0]) <= FLOAT_REL * fa)
    assert np.all(np.abs(ao[1] - ro[1]) <= FLOAT_REL * fa / cnt)
    for q, (mk, mo) in mix:
        np.testing.assert_array_equal(mk, ak)
        np.testing.assert_array_equal(mo[2], ao[2])
        assert np.all(np.abs(mo[0] - ao[0]) <= FLOAT_REL * fa)
        assert np.all(np.abs(mo[1] - ao[1]) <= FLOAT_REL * fa / cnt)


def test_processing_budget_refuses_and_releases(ops):
    """After gpu_buffer_init every device allocation is charged to the
    processing budget: one that does not fit fails with MTL_ERR_BUDGET, and
    freeing returns the charge."""
    L = abi.lib()
    used = C.c_uint64()
    abi.check(L.mtl_processing_in_use(C.byref(used)))
    base = used.value
    room = 96 << 20
    abi.check(L.mtl_init(64 << 20, base + room, 0, None, 0))
    try:
        p = C.c_void_p()
        assert L.mtl_alloc(room + (1 << 20), C.byref(p)) == abi.MTL_ERR_BUDGET
        assert "budget" in abi.last_error()
        abi.check(L.mtl_alloc(32 << 20, C.byref(p)))
        abi.check(L.mtl_processing_in_use(C.byref(used)))
        assert used.value == base + (32 << 20)
        abi.check(L.mtl_free(p))
        abi.check(L.mtl_processing_in_use(C.byref(used)))
        assert used.value == base
        assert L.mtl_free(C.c_void_p(0x1234)) == abi.MTL_ERR_INVALID_ARGUMENT
        # a group table beyond the budget: refused, nothing left charged
        n = 1000
        g = np.arange(n, dtype=np.int64) * 60_000  # a dense table of 60M x 8 B
        dg = ops.to_device(g)
        with pytest.raises(abi.MettleError) as ei:
            ops.group_agg([(dg, abi.BIGINT, 0, int(g.max()))], [(abi.AGG_COUNT, None, 0)], n)
        assert ei.value.code == abi.MTL_ERR_BUDGET
        abi.check(L.mtl_processing_in_use(C.byref(used)))
        assert used.value <= base + (64 << 20)  # at most a new scratch chunk
    finally:
        abi.check(L.mtl_shutdown())
