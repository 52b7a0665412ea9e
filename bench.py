#!/usr/bin/env python3
"""Benchmark: join -> grouped aggregate -> top-10 on one MI355X.

Workload: customers 1.5M, orders 15M, lines ~60M (seeded synthetic tables,
see mettle_amd/datagen.py); query = filter segment, o_date < cutoff <
l_ship, join lines -> orders -> customers, SUM(price * (100 - discount))
per order (exact int64), top-10 by (revenue desc, o_date asc, o_key asc).

One step = one full query over the resident columns (the star pipeline of
libmettle_hip.so).  Steps are enqueued back to back on the library stream
and bracketed by device synchronisation; value = input rows / second.
Prints ONE JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("rows/s & % HBM roofline (join+agg+top-K); queries/s + p99, batched vs serial")
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_profile(L):
    out = {}
    name = C.create_string_buffer(128)
    ms = C.c_double()
    cnt = C.c_uint64()
    i = 0
    while L.mtl_profile_get(i, name, 128, C.byref(ms), C.byref(cnt)) == 0:
        out[name.value.decode()] = (ms.value, cnt.value)
        i += 1
    return out


def measured_traffic(kernel: str, workload: str):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes recorded in profiles/traffic.json (already scaled by
    the copy-kernel calibration there), or None when not recorded for this
    kernel and workload."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if rec.get("workload") != workload:
        return None
    k = rec.get("kernels", {}).get(kernel)
    return None if k is None else k.get("hbm_bytes")


def copy_calibration(n_bytes: int = 1 << 30, reps: int = 5) -> int:
    """A plain device copy of known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters: run under rocprofv3 --pmc, each dispatch moves
    n_bytes in and n_bytes out."""
    import torch
    src = torch.empty(n_bytes, dtype=torch.uint8, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n_bytes, "reps": reps}), flush=True)
    return 0


def serving_benchmark(n_queries: int, slots: int, max_batch: int, seed: int = 20261208):
    """Serving stream: n_queries with Poisson arrivals mixing three shapes
    uniformly -- the 1M-row plain query's shape (value > literal, value*10),
    a BIGINT scan at 5-15% selectivity with a projection, and the star
    join -> SUM -> top-10 with its own segment and cutoff -- each over a
    table of log-uniform size in [1e4, 1e8] rows (prefixes of resident
    1e8-row tables), every query with its own random literals.  The same
    arrival schedule, offered at 10x the measured serial capacity, runs once
    through serial dispatch (1 slot, batches of 1: the single-query
    operators), once through continuous batching (`slots` workers, each
    taking every arrived query up to `max_batch` as one batch whose scans and
    star joins share their table passes), and once through admission alone
    (8 workers, batches of 1: overlap without shared passes); every result
    is compared between the runs.  Latency is also measured at an offered
    rate below serial capacity (0.8x), where neither queue builds up."""
    import numpy as np
    import torch
    from mettle_amd import abi, datagen, ops
    L = abi.lib()
    n_max = 100_000_000
    log("serving: generating resident tables (1e8-row scan table, star at 1e8 lines)")
    rng = np.random.default_rng(seed)
    v = rng.integers(0, 1 << 20, n_max, dtype=np.int32)
    b = rng.integers(np.iinfo(np.int64).min, np.iinfo(np.int64).max, n_max, dtype=np.int64,
                     endpoint=True)
    dv, db = ops.to_device(v), ops.to_device(b)
    del v, b
    cust, orders, lines = datagen.star_tables(scale=n_max / 60_000_000, workload=5)
    dev = {}
    for tname, cols in (("customers", cust), ("orders", orders), ("lines", lines)):
        dev[tname] = {c: ops.to_device(x) for c, x in cols.items()}
    l_order = lines["l_order"]
    n_c, n_l_all = len(cust["c_key"]), len(l_order)
    del cust, orders, lines
    torch.cuda.synchronize()
    shapes = rng.integers(0, 3, n_queries)
    sizes = np.floor(10 ** rng.uniform(4, 8, n_queries)).astype(np.int64)
    queries, bufs = [], []
    for sh, n in zip(shapes, sizes):
        if sh == 0:
            cut = int(rng.integers(0, 1 << 20))
            q = ops.scan_query([(dv, abi.INTEGER, abi.CMP_GT, cut)], int(n),
                               (dv, abi.INTEGER, abi.PROJ_MUL, 10))
            out = None
        elif sh == 1:
            cut_b = int(np.iinfo(np.int64).min + int(rng.uniform(0.05, 0.15) * (1 << 64)))
            q = ops.scan_query([(db, abi.BIGINT, abi.CMP_LT, cut_b)], int(n),
                               (db, abi.BIGINT, abi.PROJ_ADD, 7))
            out = None
        else:
            nl = int(min(n, n_l_all))
            n_o = int(l_order[nl - 1])  # orders 1..n_o cover the first nl lines
            cut = datagen.DAYS_1992_01_01 + int(rng.integers(datagen.SPAN_7Y // 4,
                                                             3 * datagen.SPAN_7Y // 4))
            d = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, n_o, nl,
                                  segment=int(rng.integers(0, 5)), cutoff=cut)
            q, out = ops.star_query(d)
        queries.append(q)
        bufs.append(out)

    def outputs(res, qs, bs):
        outs = []
        for r, q, bo in zip(res, qs, bs):
            if q.kind == abi.Q_STAR:
                outs.append(tuple((bo[i].p_row, bo[i].count, bo[i].sum_i64)
                                  for i in range(r.n_rows)))
            else:
                outs.append((r.count, r.fingerprint))
        return outs

    def run(n_slots, batch, arrivals, count=None):
        qs = queries if count is None else queries[:count]
        bs = bufs if count is None else bufs[:count]
        with ops.Server(n_slots, max_batch=batch) as srv:
            # warm-up: the largest queries, arriving together, grow the
            # workers' scratch arenas and pinned buffers before timing
            big = sorted(range(len(qs)), key=lambda i: -sizes[i])[:4 * n_slots * batch]
            t0 = L.mtl_now_ns() + 10_000_000
            for t in [srv.submit(qs[i], t0) for i in big]:
                srv.wait(t)
            t0 = L.mtl_now_ns() + 50_000_000
            tickets = [srv.submit(q, t0 + int(a)) for q, a in zip(qs, arrivals)]
            res = [srv.wait(t) for t in tickets]
        first = min(r.t_submit_ns for r in res)
        last = max(r.t_done_ns for r in res)
        lat = np.array([r.t_done_ns - r.t_submit_ns for r in res], dtype=np.float64) / 1e6
        stats = {"qps": len(res) / ((last - first) / 1e9),
                 "p50_ms": float(np.percentile(lat, 50)),
                 "p99_ms": float(np.percentile(lat, 99)), "span_s": (last - first) / 1e9,
                 "mean_batch": float(np.mean([r.batch for r in res]))}
        return stats, outputs(res, qs, bs)

    # serial capacity on a prefix, all arriving at once
    n_cal = min(500, n_queries)
    cal, _ = run(1, 1, np.zeros(n_cal), count=n_cal)
    offered = 10 * cal["qps"]
    arrivals = np.cumsum(rng.exponential(1e9 / offered, n_queries))
    log(f"serving: serial capacity ~{cal['qps']:.0f} q/s; offering {offered:.0f} q/s")
    serial, out_s = run(1, 1, arrivals)
    log(f"serving: serial {serial}")
    batched, out_b = run(slots, max_batch, arrivals)
    log(f"serving: batched ({slots} slots, batches <= {max_batch}) {batched}")
    # admission alone: 8 workers overlapping single-query runs
    shared, out_sh = run(8, 1, arrivals)
    log(f"serving: admission only (8 slots, batches of 1) {shared}")
    # latency below serial capacity: neither queue builds up
    n_low = min(2000, n_queries)
    low_rate = 0.8 * serial["qps"]
    low_arr = np.cumsum(rng.exponential(1e9 / low_rate, n_low))
    low_s, out_ls = run(1, 1, low_arr, count=n_low)
    low_b, out_lb = run(slots, max_batch, low_arr, count=n_low)
    log(f"serving at {low_rate:.0f} q/s: serial {low_s}; batched {low_b}")
    same = out_s == out_b == out_sh and out_ls == out_lb == out_s[:n_low]
    return {"queries": n_queries, "offered_qps": offered, "slots": slots,
            "max_batch": max_batch,
            "mix": "uniform over {INTEGER value > literal + value*10, BIGINT 5-15% scan + 7, "
                   "star join->SUM->top-10 (random segment, cutoff)}; rows log-uniform in "
                   "[1e4, 1e8]; random literals per query",
            "serial": serial, "batched": batched,
            "speedup": batched["qps"] / serial["qps"],
            "admission_only": dict(shared, slots=8, max_batch=1,
                                   speedup=shared["qps"] / serial["qps"]),
            "below_capacity": {"offered_qps": low_rate, "queries": n_low, "serial": low_s,
                               "batched": low_b},
            "identical_results": bool(same)}


TUNED_FILE = os.path.join(ROOT, "mettle_amd", "tuned_mi355x.txt")


def apply_tuning(args) -> dict:
    """Load the tuned launch parameters (the autotuner's saved choices for
    MI355X), or run the autotuner first with --autotune; returns the table
    and where it came from, for the JSON line."""
    from mettle_amd import abi, ops
    L = abi.lib()
    source = "defaults"
    if args.autotune:
        log("autotune:\n" + ops.autotune(ops.TUNE_ALL))
        source = "autotuned in this run"
    elif args.tuned:
        if os.path.exists(args.tuned):
            abi.check(L.mtl_tune_load(args.tuned.encode()))
            source = os.path.relpath(args.tuned, ROOT)
        else:
            log(f"no tuned parameter file at {args.tuned}: defaults")
    return {"source": source, "params": ops.tune_describe()}


def timed(fn, steps: int, warmup: int):
    """Wall time per call of fn, which runs synchronously (it returns after
    its device work is done), in ms: median over steps."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ts = []
    for _ in range(steps):
        t = time.perf_counter()
        fn()
        ts.append((time.perf_counter() - t) * 1e3)
    ts.sort()
    return ts[len(ts) // 2]


def device_timed(fn, steps: int, warmup: int) -> float:
    """Device time per call of fn, which only enqueues work on the library's
    default stream, in ms: steps calls back to back between two events on
    that stream, synchronised on both sides."""
    import torch
    from mettle_amd import abi
    L = abi.lib()
    for _ in range(warmup):
        fn()
    abi.check(L.mtl_stream_sync(None))
    torch.cuda.synchronize()
    ext = torch.cuda.ExternalStream(L.mtl_stream_native(None))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(ext)
    for _ in range(steps):
        fn()
    e1.record(ext)
    abi.check(L.mtl_stream_sync(None))
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def scan_workload(args) -> dict:
    """100M-row scan-filter-project: predicates at 1/10/50% selectivity on
    INTEGER and BIGINT columns with a projection, the whole operator chain
    (filter + compaction + projection, fused in mtl_filter_project) per
    call: `ms` is device time per call with calls back to back, `latency_ms`
    the wall time of one synchronised call; mtl_filter then mtl_project is
    timed beside it; algorithmic bytes = N x predicate width + S x (row id 4
    + output width)."""
    import numpy as np
    import torch
    from mettle_amd import abi, datagen, ops
    n = 100_000_000
    log("scan: generating 100M rows")
    t = datagen.scan_table(n)
    dev = {c: ops.to_device(t[c]) for c in ("c_i", "c_b")}
    del t
    sel = torch.empty(n, dtype=torch.int32, device="cuda")
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    ovf = torch.zeros(1, dtype=torch.int32, device="cuda")
    L = abi.lib()
    rows = []
    for cname, ctype, width in (("c_i", abi.INTEGER, 4), ("c_b", abi.BIGINT, 8)):
        lo = -(1 << 31) if width == 4 else -(1 << 63)
        span = 1 << 32 if width == 4 else 1 << 64
        for frac in (0.01, 0.10, 0.50):
            lit = lo + int(frac * span)
            preds = (abi.Pred * 1)(abi.make_pred(dev[cname].data_ptr(), ctype, abi.CMP_LT, lit))

            def fused_async():
                abi.check(L.mtl_filter_project(preds, 1, n, None, None, 0,
                                               dev[cname].data_ptr(), ctype, abi.PROJ_ADD, 1,
                                               sel.data_ptr(), cnt.data_ptr(), out.data_ptr(),
                                               ovf.data_ptr(), None))

            def fused():
                fused_async()
                abi.check(L.mtl_stream_sync(None))

            def separate():
                abi.check(L.mtl_filter(preds, 1, n, None, None, 0, sel.data_ptr(),
                                       cnt.data_ptr(), None))
                abi.check(L.mtl_project(dev[cname].data_ptr(), ctype, sel.data_ptr(),
                                        cnt.data_ptr(), n, abi.PROJ_ADD, 1, out.data_ptr(),
                                        ovf.data_ptr(), None))
                abi.check(L.mtl_stream_sync(None))
            def order_by():
                got = C.c_uint32()
                abi.check(L.mtl_sort(dev[cname].data_ptr(), ctype, sel.data_ptr(),
                                     cnt.data_ptr(), n, 0, perm.data_ptr(), C.byref(got), None))
                abi.check(L.mtl_stream_sync(None))
            ms_sep = timed(separate, args.steps, args.warmup)
            ms_lat = timed(fused, args.steps, args.warmup)
            # throughput: calls back to back, inputs resident, one
            # synchronisation on each side (the latency form above adds the
            # host's launch and synchronisation to every call)
            ms = device_timed(fused_async, args.steps, args.warmup)
            s_rows = int(cnt.item())
            # ORDER BY the filtered column over the selection (the config-2
            # query with its optional ORDER BY): radix sort of S (key, row)
            # pairs; algorithmic bytes = S x (key width + row id in + out)
            ms_sort = timed(order_by, max(3, args.steps // 3), 1)
            sort_alg = s_rows * (width + 8)
            alg = n * width + s_rows * (4 + width)
            rows.append({"column": cname, "selectivity": frac, "selected": s_rows, "ms": ms,
                         "latency_ms": ms_lat, "ms_filter_then_project": ms_sep,
                         "order_by": {"ms": ms_sort,
                                      "rows_per_s": s_rows / (ms_sort * 1e-3),
                                      "algorithmic_GBs": sort_alg / (ms_sort * 1e-3) / 1e9},
                         "rows_per_s": n / (ms * 1e-3), "algorithmic_bytes": alg,
                         "GBs": alg / (ms * 1e-3) / 1e9,
                         "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS})
            log(f"scan {cname} {frac:.0%}: fused {ms:.3f} ms ({rows[-1]['GBs']:.0f} GB/s; "
                f"{ms_lat:.3f} ms synchronised per call), "
                f"filter then project {ms_sep:.3f} ms, ORDER BY {ms_sort:.3f} ms")
    # the scalar hooks (gpu_add_scalar, gpu_lt_scalar) on one DuckDB vector
    # of 2048 INTEGER values: host vector in, result out, per call
    vec = np.arange(2048, dtype=np.int32)
    hout = np.empty(2048, dtype=np.int32)
    hooks = {}
    for name, hook in (("add", abi.HOOK_ADD), ("lt", abi.HOOK_LT)):
        def hook_call():
            abi.check(L.mtl_scalar_hook(hook, abi.INTEGER, vec.ctypes.data, len(vec),
                                        abi.scalar_bits(7, abi.INTEGER), hout.ctypes.data))
        hooks[name] = {"us_per_vector": timed(hook_call, 200, 20) * 1e3, "values": len(vec)}
    log(f"scalar hooks on 2048 values: add {hooks['add']['us_per_vector']:.1f} us, "
        f"lt {hooks['lt']['us_per_vector']:.1f} us")
    return {"metric": "rows/s & % HBM roofline (scan-filter-project)",
            "value": min(r["rows_per_s"] for r in rows), "unit": "rows/s", "hooks": hooks,
            "config": {"workload": "scan-filter-project 100M rows, "
                                   "col < literal + projection col + 1", "rows": n},
            "variants": rows, "dtype": "int32/int64", "data": "synthetic"}


def group_workload(args, dense: bool) -> dict:
    """Skewed grouped aggregate: 200M rows, g BIGINT Zipf(1.1) over 10M keys
    (or 1000 uniform keys), SUM/COUNT/MIN/MAX(v_i), SUM/AVG(v_f), then
    top-100 by SUM(v_i); algorithmic bytes = 200M x (8 + 8 + 4)."""
    import ctypes as C
    import torch
    from mettle_amd import abi, datagen, ops
    n = 200_000_000
    log("group: generating 200M rows")
    t = datagen.group_table(n, dense_keys=1000 if dense else None)
    gmax = int(t["g"].max())
    dg, dvi, dvf = (ops.to_device(t[c]) for c in ("g", "v_i", "v_f"))
    del t
    L = abi.lib()
    aggs = [(abi.AGG_SUM, dvi, abi.BIGINT), (abi.AGG_COUNT, None, 0),
            (abi.AGG_MIN, dvi, abi.BIGINT), (abi.AGG_MAX, dvi, abi.BIGINT),
            (abi.AGG_SUM, dvf, abi.FLOAT), (abi.AGG_AVG, dvf, abi.FLOAT)]
    # the C ABI directly (as the join and scan workloads): one descriptor and
    # the output buffers made once, then group-by + top-100 per call
    import numpy as np
    desc = abi.make_group_desc(n, [(dg.data_ptr(), abi.BIGINT, 0, gmax)],
                               [(f, c.data_ptr() if c is not None else None, t)
                                for f, c, t in aggs])
    keys_out = np.empty((100, 1), dtype=np.int64)
    outs = [np.empty(100, dtype=abi.agg_result_dtype(f, t)) for f, _, t in aggs]
    ptrs = (C.c_void_p * len(outs))(*[o.ctypes.data for o in outs])
    got = C.c_uint32()
    res = {}

    def call():
        h = C.c_void_p()
        abi.check(L.mtl_group_agg(C.byref(desc), C.byref(h), None))
        try:
            abi.check(L.mtl_groups_topk(h, 0, 1, 100, keys_out.ctypes.data, ptrs,
                                        C.byref(got)))
        finally:
            L.mtl_groups_release(h)
        res["topk"] = got.value
    ms = timed(call, args.steps, args.warmup)
    alg = n * 20
    gbs = alg / (ms * 1e-3) / 1e9
    log(f"group ({'1000 keys' if dense else 'Zipf 10M keys'}): {ms:.3f} ms, {gbs:.0f} GB/s")
    return {"metric": "rows/s & % HBM roofline (grouped aggregate + top-100)",
            "value": n / (ms * 1e-3), "unit": "rows/s", "ms_per_step": ms,
            "config": {"workload": "grouped aggregate 200M rows, "
                                   + ("g uniform over 1000 keys" if dense else
                                      "g Zipf(1.1) over 10M keys"), "rows": n},
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "algorithmic_bytes": alg},
            "dtype": "int64/float32 (exact 128-bit integer sums, float64 float sums)",
            "data": "synthetic"}


def cpu_baseline(scale: float, seconds: float):
    """Oracle (C restatement, 1 thread) on a bounded sample of the workload;
    also checks the GPU result on the same sample."""
    from mettle_amd import abi, datagen, ops
    from oracle import oracle as orc
    tables = datagen.star_tables(scale=scale)
    n_c, n_o, n_l = (len(tables[0]["c_key"]), len(tables[1]["o_key"]),
                     len(tables[2]["l_order"]))
    by = dict(zip(["customers", "orders", "lines"], tables))
    hd = datagen.star_desc(lambda t, c: by[t][c].ctypes.data, n_c, n_o, n_l)
    runs, t0 = 0, time.perf_counter()
    while True:
        st, ref = orc.star_topk(hd)
        runs += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    rows = n_c + n_o + n_l
    dev = {t: {c: ops.to_device(v) for c, v in cols.items()} for t, cols in by.items()}
    dd = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, n_o, n_l)
    got = ops.star_topk(dd)
    same = (st == abi.MTL_OK and [(r.p_row, r.count, r.sum_i64) for r in got] ==
            [(r.p_row, r.count, r.sum_i64) for r in ref])
    return {"value": rows * runs / el, "unit": "rows/s", "cores": 1, "kind": "port",
            "sample": (f"oracle/mettle_oracle.c star pipeline, scale {scale} "
                       f"({n_c}+{n_o}+{n_l} rows), {runs} runs in {el:.1f}s, 1 thread"),
            "gpu_matches_oracle_on_sample": bool(same)}


def join_workload(args) -> dict:
    """PK/FK join at the config-3 sizes with keys too sparse for the dense
    table: build on 15M unique BIGINT keys, probe with 60M foreign keys
    (every probe matches).  Times the global hash table (MTL_JOIN_HASH,
    pairs in probe order) and the LDS-staged radix partitions
    (MTL_JOIN_RADIX), build + probe per call; algorithmic bytes = build keys
    8 B x 15M + probe keys 8 B x 60M + pairs 8 B x 60M."""
    import numpy as np
    import torch
    from mettle_amd import abi
    nb, npr = 15_000_000, 60_000_005
    rng = np.random.default_rng(3)
    log("join: generating keys")
    bk = torch.from_numpy((rng.permutation(nb).astype(np.int64) * 2654435761 - (1 << 50))).cuda()
    pk = bk[torch.randint(0, nb, (npr,), device="cuda")].contiguous()
    out_p = torch.empty(npr, dtype=torch.int32, device="cuda")
    out_b = torch.empty(npr, dtype=torch.int32, device="cuda")
    cnt = torch.zeros(4, dtype=torch.int32, device="cuda")
    dup = torch.zeros(4, dtype=torch.int32, device="cuda")
    L = abi.lib()
    alg = nb * 8 + npr * 8 + npr * 8
    rows = []
    for strat, name in ((abi.JOIN_HASH, "hash"), (abi.JOIN_RADIX, "radix")):
        table = torch.empty(int(L.mtl_join_table_bytes(strat, nb, 0)), dtype=torch.uint8,
                            device="cuda")

        def call():
            abi.check(L.mtl_join_build(strat, bk.data_ptr(), abi.BIGINT, None, None, nb, 0, 0,
                                       table.data_ptr(), table.numel(), dup.data_ptr(), None))
            abi.check(L.mtl_join_probe(strat, table.data_ptr(), table.numel(), 0, nb,
                                       pk.data_ptr(), abi.BIGINT, None, None, npr,
                                       out_p.data_ptr(), out_b.data_ptr(), cnt.data_ptr(), None))
            abi.check(L.mtl_stream_sync(None))
        ms = timed(call, args.steps, args.warmup)
        L.mtl_profile_reset()
        L.mtl_profile_enable(1)
        call()
        L.mtl_profile_enable(0)
        kp = kernel_profile(L)
        L.mtl_profile_reset()
        assert int(cnt[0].item()) == npr and int(dup[0].item()) == 0
        gbs = alg / (ms * 1e-3) / 1e9
        rows.append({"strategy": name, "ms": ms, "algorithmic_GBs": gbs,
                     "frac": gbs / HBM_PEAK_GBS,
                     "kernels_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in kp.items()}})
        log(f"join {name}: {ms:.3f} ms ({gbs:.0f} GB/s algorithmic)")
    best = min(rows, key=lambda r: r["ms"])
    return {"metric": "probe rows/s & % HBM roofline (PK/FK join, sparse keys)",
            "value": npr / (best["ms"] * 1e-3), "unit": "rows/s", "ms_per_step": best["ms"],
            "config": {"workload": "join build 15M unique BIGINT keys, probe 60M",
                       "build_rows": nb, "probe_rows": npr},
            "roofline": {"bound": "hbm", "achieved": best["algorithmic_GBs"],
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": best["frac"],
                         "algorithmic_bytes": alg},
            "variants": rows, "dtype": "int64", "data": "synthetic"}


def ingest_workload(args) -> dict:
    """Ingestion at the config-3 size (the pinned.rs replacement): the four
    columns of lines (60M rows: l_order and l_ship 4 B, price and discount
    8 B; 1.44 GB) uploaded through the HBM column cache (mtl_cache_put:
    pageable host memory -> two pinned chunks -> device, a new version each
    call so every call uploads), then the same table registered from Arrow
    (mtl_catalog_ingest_arrow: per-column host conversion + the same
    uploads + catalog).  Inputs start in host memory: this is the
    PCIe-inclusive path, not an HBM-resident rate."""
    import ctypes as C
    import numpy as np
    import pyarrow as pa
    from mettle_amd import abi, datagen
    from mettle_amd import sql as S
    L = abi.lib()
    abi.check(L.mtl_init(8 << 30, 8 << 30, 0, None, 0))
    _, _, lines = datagen.star_tables()
    cols = [np.ascontiguousarray(lines[c]) for c in ("l_order", "l_ship", "price", "discount")]
    n = len(cols[0])
    nbytes = sum(c.nbytes for c in cols)
    ver = [0]

    def put():
        ver[0] += 1
        for i, c in enumerate(cols):
            p = C.c_void_p()
            abi.check(L.mtl_cache_put(77, i, c.ctypes.data, c.nbytes, ver[0], C.byref(p)))
            abi.check(L.mtl_cache_release(77, i))
        abi.check(L.mtl_stream_sync(None))
    ms_put = timed(put, max(3, args.steps // 5), 1)
    t = pa.table({"l_order": cols[0], "l_ship": pa.array(cols[1], type=pa.date32()),
                  "price": cols[2], "discount": cols[3]})
    ms_ing = timed(lambda: S.ingest_arrow("lines_ingest", t), 3, 1)
    S.drop("lines_ingest")
    abi.check(L.mtl_cache_reset())
    abi.check(L.mtl_shutdown())
    up = nbytes / (ms_put * 1e-3) / 1e9
    ing = nbytes / (ms_ing * 1e-3) / 1e9
    log(f"ingest: cache upload {ms_put:.1f} ms ({up:.1f} GB/s), Arrow ingestion "
        f"{ms_ing:.1f} ms ({ing:.1f} GB/s)")
    return {"metric": "rows/s & GB/s (host -> HBM column cache, Arrow ingestion)",
            "value": n / (ms_ing * 1e-3), "unit": "rows/s", "ms_per_step": ms_ing,
            "config": {"workload": "ingest lines 60M rows x 4 columns (1.44 GB)", "rows": n,
                       "bytes": nbytes},
            "cache_upload": {"ms": ms_put, "GBs": up},
            "arrow_ingest": {"ms": ms_ing, "GBs": ing},
            "dtype": "int32/int64", "data": "synthetic"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-scale", type=float, default=0.1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-queries", type=int, default=10000)
    ap.add_argument("--serve-slots", type=int, default=1)
    ap.add_argument("--serve-batch", type=int, default=256)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--workload", choices=["star", "scan", "group", "group-dense", "join",
                                           "ingest"],
                    default="star", help="star: the headline line; the others measure "
                    "configs 2 and 4 (parity configs, reported separately)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--tuned", default=TUNED_FILE,
                    help="tuned launch parameters to load ('' keeps the defaults)")
    ap.add_argument("--autotune", action="store_true",
                    help="run the autotuner on this device first (report on stderr)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    tuned = apply_tuning(args)
    if args.workload != "star":
        if args.gpus != 1:
            print(json.dumps({"error": "this library runs on one GPU"}))
            return 2
        if args.workload == "scan":
            r = scan_workload(args)
        elif args.workload == "join":
            r = join_workload(args)
        elif args.workload == "ingest":
            r = ingest_workload(args)
        else:
            r = group_workload(args, args.workload == "group-dense")
        r.update({"n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                  "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                  "tuned": tuned})
        print(json.dumps(r), flush=True)
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is designed but not run)"}))
        return 2

    import torch
    from mettle_amd import abi, datagen, ops

    L = abi.lib()
    log("generating tables")
    t0 = time.time()
    cust, orders, lines = datagen.star_tables(scale=args.scale)
    n_c, n_o, n_l = len(cust["c_key"]), len(orders["o_key"]), len(lines["l_order"])
    log(f"tables: customers {n_c}, orders {n_o}, lines {n_l} ({time.time() - t0:.1f}s)")
    dev = {}
    for tname, cols in (("customers", cust), ("orders", orders), ("lines", lines)):
        dev[tname] = {c: ops.to_device(v) for c, v in cols.items()}
    torch.cuda.synchronize()
    log("columns resident in HBM")
    desc = datagen.star_desc(lambda t, c: dev[t][c].data_ptr(), n_c, n_o, n_l)

    s = C.c_void_p()
    abi.check(L.mtl_stream_create(C.byref(s)))
    out_dev = torch.empty(desc.k * C.sizeof(abi.StarRow) * (args.steps + args.warmup + 1),
                          dtype=torch.uint8, device="cuda")
    out_n = torch.zeros(2 * (args.steps + args.warmup + 1), dtype=torch.int32, device="cuda")
    row_bytes = desc.k * C.sizeof(abi.StarRow)

    def step(i):
        abi.check(L.mtl_star_topk_async(C.byref(desc), out_dev.data_ptr() + i * row_bytes,
                                        out_n.data_ptr() + 8 * i, s))

    # correctness of one synchronous run (result also checked vs the oracle
    # in the cpu_baseline leg on a sample)
    rows = (abi.StarRow * desc.k)()
    nrows = C.c_uint32()
    abi.check(L.mtl_star_topk(C.byref(desc), rows, C.byref(nrows), s))
    log(f"top row: order {int(orders['o_key'][rows[0].p_row])} revenue {rows[0].sum_i64}")

    for i in range(args.warmup):
        step(i)
    abi.check(L.mtl_stream_sync(s))
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} done; timing {args.steps} steps")

    ext = torch.cuda.ExternalStream(L.mtl_stream_native(s))
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t_wall = time.perf_counter()
    e0.record(ext)
    for i in range(args.steps):
        step(args.warmup + i)
    e1.record(ext)
    abi.check(L.mtl_stream_sync(s))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall
    dev_ms = e0.elapsed_time(e1)
    err_bits = out_n.cpu().numpy()[1::2]
    if (err_bits != 0).any():
        abi.check(L.mtl_star_status(int(err_bits[err_bits != 0][0])))
    ms_per_step = dev_ms / args.steps
    in_rows = n_c + n_o + n_l
    value = in_rows * args.steps / (dev_ms / 1e3)
    alg_bytes = datagen.star_algorithmic_bytes(n_c, n_o, n_l)
    log(f"{ms_per_step:.4f} ms/step device, wall {wall / args.steps * 1e3:.4f} ms/step")

    # serial latency: synchronous single queries (host sync + 10-row download)
    lat = []
    for _ in range(min(args.steps, 30)):
        t = time.perf_counter()
        abi.check(L.mtl_star_topk(C.byref(desc), rows, C.byref(nrows), s))
        lat.append((time.perf_counter() - t) * 1e3)
    lat.sort()

    # per-kernel durations (hipEvents on the library stream)
    L.mtl_profile_reset()
    L.mtl_profile_enable(1)
    prof_steps = min(args.steps, 20)
    for i in range(prof_steps):
        step(i)
    abi.check(L.mtl_stream_sync(s))
    L.mtl_profile_enable(0)
    prof = kernel_profile(L)
    per_kernel = {k: round(ms / max(c, 1) * 1e3, 2) for k, (ms, c) in prof.items()}
    dom = max(prof, key=lambda k: prof[k][0])
    dom_us = prof[dom][0] / prof[dom][1] * 1e3
    counts = datagen.star_counts(cust, orders, lines)
    kbytes = datagen.star_kernel_bytes(n_c, n_o, n_l, counts)
    dom_bytes = kbytes.get(dom, 0)
    achieved = dom_bytes / (dom_us * 1e-6) / 1e9
    kernels = {k: {"us": per_kernel[k], "bytes": kbytes.get(k, 0),
                   "GBs": round(kbytes.get(k, 0) / (per_kernel[k] * 1e-6) / 1e9, 1)
                   if per_kernel[k] > 0 else None}
               for k in per_kernel}
    log(f"kernels: {kernels}")
    traffic = measured_traffic(dom, f"lines {n_l}")

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic (numpy PCG64 seed 20261203+3), columns resident in HBM",
        "config": {"workload": "join->group-agg->top-10: customers 1.5M, orders 15M, "
                               f"lines {n_l}",
                   "rows": in_rows, "scale": args.scale, "k": 10},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic, "algorithmic_bytes": dom_bytes,
                     "avg_launch_us": dom_us},
        "pipeline": {"algorithmic_bytes": alg_bytes, "device_ms": ms_per_step,
                     "achieved_GBs": alg_bytes / (ms_per_step * 1e-3) / 1e9,
                     "frac": alg_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "kernels": kernels, "stage_rows": counts},
        "latency_ms": {"serial_p50": lat[len(lat) // 2], "serial_p99": lat[-1],
                       "queries_per_s_serial": 1e3 / (sum(lat) / len(lat)),
                       "queries_per_s_batched": args.steps / (dev_ms / 1e3)},
        "tuned": tuned,
    }
    L.mtl_stream_destroy(s)
    del dev
    if not args.no_serve:
        result["serving"] = serving_benchmark(args.serve_queries, args.serve_slots,
                                              args.serve_batch)
        result["latency_ms"]["queries_per_s_serial"] = result["serving"]["serial"]["qps"]
        result["latency_ms"]["queries_per_s_batched"] = result["serving"]["batched"]["qps"]
        result["latency_ms"]["serving_p50_ms"] = result["serving"]["batched"]["p50_ms"]
        result["latency_ms"]["serving_p99_ms"] = result["serving"]["batched"]["p99_ms"]
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle, 1 thread)")
        result["cpu_baseline"] = cpu_baseline(args.cpu_scale, args.cpu_seconds)
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
