"""Benchmark of the MI355X hash groupby-aggregate.

Default workload (the headline): 100M rows, int64 key drawn from 1M distinct
values spread over the whole int64 range; value columns int32 uniform in
[-1e6, 1e6], float64 N(0,1) and float32 N(0,1), each 5% null; sum, count,
min, max and mean of every value column. One step = one
cudf::groupby::groupby::aggregate call through the C ABI on device-resident
inputs (output allocation included).

Prints one JSON line. `value` is rows/s over the timed steps; `roofline`
prices the whole call (input columns read once and results written once, over
the measured time per step) and, in `dominant_kernel`, the longest kernel with
its algorithmic bytes over its own average duration, measured with HIP events
on its launch stream in a separate pass; `cpu_baseline` times pandas (what
the reference runs on the CPU: the cudf.pandas fallback) on a bounded sample of
the same data on this host, with the oracle's C port beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "groupby rows/s & % HBM peak (100M rows); batched q/s, p50/p99 at fixed QPS"


def log(msg: str) -> None:
    """Progress on stderr (stdout carries only the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
HBM_PEAK_GBS = 8000.0

# library profiling scopes (mgb_profiling_read) and the kernels they time
PROFILED = ("aggregate", "lds_aggregate", "fill", "finalize", "partition_hist", "partition_scatter", "estimate",
            "pass2", "sort_aggs")
KERNEL_SYMBOL = {"aggregate": "mgb::aggregate_kernel", "lds_aggregate": "mgb::lds_agg_kernel<false>",
                 "partition_scatter": "mgb::part_scatter_tiled_kernel", "partition_hist": "mgb::part_hist_kernel",
                 "pass2": "mgb::pass2_kernel", "sort_aggs": "sort-path kernels and hipCUB radix sorts"}
# kernels of the compact partitioned execution (single integer key, up to three
# 4/8-byte value columns), compiled per shape: a scope's kernel is found by
# prefix in the traffic summary
COMPACT_PREFIX = {"lds_aggregate": "mgb::lean_agg_kernel<", "partition_scatter": "mgb::lean_scatter_kernel<"}


def kernel_name(raw: str) -> str:
    """A rocprofv3 kernel name without 'void', arguments or spaces."""
    return raw.split("(")[0].removeprefix("void ").replace(" ", "")


def pack_bits(valid):
    """bool tensor [n] -> int32 bitmask words (LSB first), padded to 64 bytes."""
    import torch

    n = valid.numel()
    words = (n + 31) // 32
    alloc_words = ((words * 4 + 63) // 64 * 64) // 4
    padded = torch.zeros(words * 32, dtype=torch.int64, device=valid.device)
    padded[:n] = valid.to(torch.int64)
    weights = (torch.ones(32, dtype=torch.int64, device=valid.device) << torch.arange(32, device=valid.device))
    w = (padded.view(words, 32) * weights).sum(dim=1)
    w = torch.where(w >= 2 ** 31, w - 2 ** 32, w).to(torch.int32)
    out = torch.zeros(alloc_words, dtype=torch.int32, device=valid.device)
    out[:words] = w
    return out


def make_c2(rows: int, distinct: int, seed: int, device: str = "cuda"):
    """Seeded config-2 data generated on the device."""
    import torch

    import cudf_amd as ca

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    uniq = torch.unique(torch.randint(-2 ** 63, 2 ** 63 - 1, (distinct,), generator=g, device=device,
                                      dtype=torch.int64))
    keys = uniq[torch.randint(0, uniq.numel(), (rows,), generator=g, device=device)]
    v_i32 = torch.randint(-1_000_000, 1_000_001, (rows,), generator=g, device=device, dtype=torch.int32)
    v_f64 = torch.randn(rows, generator=g, device=device, dtype=torch.float64)
    v_f32 = torch.randn(rows, generator=g, device=device, dtype=torch.float32)
    cols = []
    for t, tid in ((v_i32, ca.TypeId.INT32), (v_f64, ca.TypeId.FLOAT64), (v_f32, ca.TypeId.FLOAT32)):
        valid = torch.rand(rows, generator=g, device=device) >= 0.05
        mask = pack_bits(valid)
        nulls = int((~valid).sum().item())
        del valid
        cols.append(ca.Column.from_torch(t, tid, mask, nulls))
    key_col = ca.Column.from_torch(keys, ca.TypeId.INT64)
    return key_col, cols, int(uniq.numel())


def c2_requests(cols):
    import cudf_amd as ca
    from cudf_amd import aggregation as agg

    return [ca.GroupByRequest(c, [agg.sum(), agg.count(), agg.min(), agg.max(), agg.mean()]) for c in cols]


def zipf_ranks(n_rows: int, n_keys: int, s: float, g, device: str = "cuda"):
    """Ranks in [0, n_keys) with P(rank r) ~ (r+1)^-s, by inverting the
    continuous approximation of the Zipf CDF."""
    import torch

    u = torch.rand(n_rows, generator=g, device=device, dtype=torch.float64)
    a = 1.0 - s
    x = (1.0 + u * (float(n_keys) ** a - 1.0)) ** (1.0 / a)
    return torch.clamp(x.floor().to(torch.int64) - 1, 0, n_keys - 1)


def make_workload(name: str, rows: int, seed: int, device: str = "cuda", distinct: int = 1_000_000):
    """(key columns, requests, input bytes per row, output bytes per group, description)."""
    import torch

    import cudf_amd as ca
    from cudf_amd import aggregation as agg

    g = torch.Generator(device=device)
    g.manual_seed(seed)
    if name == "c2":
        key_col, cols, distinct = make_c2(rows, distinct, seed, device)
        return ([key_col], c2_requests(cols), 8 + 4 + 8 + 4 + 3 / 8, 96 + 12 / 8,
                f"groupby-aggregate, {rows:,} rows, int64 key over {distinct:,} distinct values, int32/f64/f32 "
                "values 5% null, sum/count/min/max/mean per value column")
    if name == "c2v":
        # the headline data with the second-pass aggregations on every value column
        key_col, cols, distinct = make_c2(rows, distinct, seed, device)
        req = [ca.GroupByRequest(c, [agg.mean(), agg.variance(), agg.std(), agg.argmin(), agg.argmax()])
               for c in cols]
        return ([key_col], req, 8 + 4 + 8 + 4 + 3 / 8, 8 + 3 * (8 + 8 + 8 + 4 + 4) + 15 / 8,
                f"groupby-aggregate, {rows:,} rows, int64 key over {distinct:,} distinct values, int32/f64/f32 "
                "values 5% null, mean/variance/std/argmin/argmax per value column")
    if name == "c2s":
        # the headline data with aggregations cudf serves from its sort path
        key_col, cols, distinct = make_c2(rows, distinct, seed, device)
        req = [ca.GroupByRequest(c, [agg.median(), agg.quantile(0.9), agg.nunique(), agg.nth_element(0)])
               for c in cols]
        return ([key_col], req, 8 + 4 + 8 + 4 + 3 / 8, 8 + 3 * (8 + 8 + 4) + 4 + 8 + 4 + 15 / 8,
                f"groupby-aggregate, {rows:,} rows, int64 key over {distinct:,} distinct values, int32/f64/f32 "
                "values 5% null, median/quantile(0.9)/nunique/nth_element(0) per value column")
    if name == "c1":
        keys = torch.randint(0, 1000, (rows,), generator=g, device=device, dtype=torch.int32)
        vals = torch.rand(rows, generator=g, device=device, dtype=torch.float64)
        req = [ca.GroupByRequest(ca.Column.from_torch(vals, ca.TypeId.FLOAT64),
                                 [agg.sum(), agg.count(), agg.mean(), agg.min(), agg.max()])]
        return ([ca.Column.from_torch(keys, ca.TypeId.INT32)], req, 12, 4 + 8 + 4 + 8 + 8 + 8,
                f"groupby-aggregate, {rows:,} rows, int32 key in [0,1000), f64 value in [0,1), "
                "sum/count/mean/min/max")
    if name == "c3a":
        n_keys = 50_000_000
        ka = torch.randint(-2 ** 31, 2 ** 31 - 1, (n_keys,), generator=g, device=device, dtype=torch.int32)
        kb = torch.randint(-2 ** 63, 2 ** 63 - 1, (n_keys,), generator=g, device=device, dtype=torch.int64)
        perm = torch.randperm(n_keys, generator=g, device=device)
        idx = perm[zipf_ranks(rows, n_keys, 1.1, g, device)]
        del perm
        a, b = ka[idx], kb[idx]
        del idx, ka, kb
        vals = torch.randn(rows, generator=g, device=device, dtype=torch.float64)
        req = [ca.GroupByRequest(ca.Column.from_torch(vals, ca.TypeId.FLOAT64), [agg.sum(), agg.count(), agg.max()])]
        return ([ca.Column.from_torch(a, ca.TypeId.INT32), ca.Column.from_torch(b, ca.TypeId.INT64)], req, 20,
                4 + 8 + 8 + 4 + 8,
                f"groupby-aggregate, {rows:,} rows, composite (int32, int64) key Zipf s=1.1 over 50M distinct tuples, "
                "f64 value, sum/count/max")
    if name == "c3b":
        uniq = torch.randint(-2 ** 31, 2 ** 31 - 1, (16,), generator=g, device=device, dtype=torch.int32)
        keys = uniq[torch.randint(0, 16, (rows,), generator=g, device=device)]
        vals = torch.randn(rows, generator=g, device=device, dtype=torch.float64)
        req = [ca.GroupByRequest(ca.Column.from_torch(vals, ca.TypeId.FLOAT64), [agg.sum(), agg.count(), agg.max()])]
        return ([ca.Column.from_torch(keys, ca.TypeId.INT32)], req, 12, 4 + 8 + 4 + 8,
                f"groupby-aggregate, {rows:,} rows, int32 key with 16 distinct values, f64 value, sum/count/max")
    raise SystemExit(f"unknown workload {name}")


DEFAULT_ROWS = {"c1": 1_000_000, "c2": 100_000_000, "c2v": 100_000_000, "c2s": 100_000_000, "c3a": 200_000_000, "c3b": 200_000_000}


def make_c4(num_queries: int, seed: int, device: str = "cuda", max_rows: float = 1e7):
    """Batched workload: per query rows log-uniform in [1e3, max_rows], groups
    log-uniform in [10, 1e5] (at most the rows), int32 keys from that many
    distinct values, float64 N(0,1) values; sum/count/mean."""
    import numpy as np
    import torch

    import cudf_amd as ca
    from cudf_amd import aggregation as agg

    rng = np.random.default_rng(seed)
    rows = np.rint(10 ** rng.uniform(3, np.log10(max_rows), num_queries)).astype(np.int64)
    card = np.minimum(rows, np.rint(10 ** rng.uniform(1, 5, num_queries))).astype(np.int64)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    queries = []
    for q in range(num_queries):
        idx = torch.randint(0, int(card[q]), (int(rows[q]),), generator=g, device=device, dtype=torch.int64)
        a = int(rng.integers(1, 2 ** 31)) * 2 + 1       # odd: a bijection of the 32-bit integers
        b = int(rng.integers(0, 2 ** 32))
        k = (idx * a + b) % (2 ** 32)
        keys = torch.where(k >= 2 ** 31, k - 2 ** 32, k).to(torch.int32)
        del idx, k
        vals = torch.randn(int(rows[q]), generator=g, device=device, dtype=torch.float64)
        queries.append((ca.Table([ca.Column.from_torch(keys, ca.TypeId.INT32)]),
                        [ca.GroupByRequest(ca.Column.from_torch(vals, ca.TypeId.FLOAT64),
                                           [agg.sum(), agg.count(), agg.mean()])]))
    return queries, rows, card


def make_c5_tables(num_tables: int, seed: int, device: str = "cuda", max_rows: float = 5e7):
    """Serving tables: rows log-uniform in [1e4, max_rows], groups log-uniform in
    [10, 1e6] (at most the rows), int64 key, float64 N(0,1) and int32 uniform
    [-1e6, 1e6] values, each 2% null."""
    import numpy as np
    import torch

    import cudf_amd as ca

    rng = np.random.default_rng(seed)
    rows = np.rint(10 ** rng.uniform(4, np.log10(max_rows), num_tables)).astype(np.int64)
    card = np.minimum(rows, np.rint(10 ** rng.uniform(1, 6, num_tables))).astype(np.int64)
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    tables = []
    for t in range(num_tables):
        n = int(rows[t])
        pool = torch.randint(-2 ** 63, 2 ** 63 - 1, (int(card[t]),), generator=g, device=device, dtype=torch.int64)
        keys = pool[torch.randint(0, int(card[t]), (n,), generator=g, device=device)]
        f64 = torch.randn(n, generator=g, device=device, dtype=torch.float64)
        i32 = torch.randint(-1_000_000, 1_000_001, (n,), generator=g, device=device, dtype=torch.int32)
        cols = [ca.Column.from_torch(keys, ca.TypeId.INT64)]
        for v, tid in ((f64, ca.TypeId.FLOAT64), (i32, ca.TypeId.INT32)):
            valid = torch.rand(n, generator=g, device=device) >= 0.02
            cols.append(ca.Column.from_torch(v, tid, pack_bits(valid), int((~valid).sum().item())))
        tables.append(cols)
    return tables, rows, card


def serve_run(tables, rate: float, n_queries: int, batching: bool, seed: int, budget: int, reserve: int = 0,
              max_inflight: int = 2, max_batch: int = 256):
    """Poisson arrivals at `rate` queries/s on random tables; returns latencies and throughput."""
    import threading

    import numpy as np

    import cudf_amd as ca
    from cudf_amd import aggregation as agg

    aggs = [agg.sum(), agg.count(), agg.min(), agg.max(), agg.mean()]
    srv = ca.Server(hbm_budget_bytes=budget, max_inflight=max_inflight, max_batch_queries=max_batch,
                    batching=batching, reserve_bytes=reserve)
    tids = [srv.add_table(t) for t in tables]
    # one query per table first, not timed: the steady state of a server
    # whose tables are resident (workspaces sized, code loaded)
    warm = [srv.submit(t, [0], [(1, aggs), (2, aggs)]) for t in tids]
    for tk in warm:
        srv.wait(tk, 600_000)
        srv.release(tk)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_queries)
    picks = rng.integers(0, len(tids), n_queries)
    tickets = [None] * n_queries
    lat = [0.0] * n_queries
    done_t = [0.0] * n_queries
    submitted = threading.Semaphore(0)

    def collect():
        for i in range(n_queries):
            submitted.acquire()
            _, lat[i] = srv.wait(tickets[i], 600_000)
            done_t[i] = time.perf_counter()
            srv.release(tickets[i])

    th = threading.Thread(target=collect)
    th.start()
    t0 = time.perf_counter()
    due = t0
    for i in range(n_queries):
        due += gaps[i]
        while True:
            now = time.perf_counter()
            if now >= due:
                break
            if due - now > 0.0005:
                time.sleep(due - now - 0.0003)
        tickets[i] = srv.submit(tids[int(picks[i])], [0], [(1, aggs), (2, aggs)])
        submitted.release()
    t_submitted = time.perf_counter() - t0
    th.join()
    elapsed = max(done_t) - t0
    st = srv.stats()
    srv.close()
    lat_ms = np.asarray(lat)
    return {
        "offered_qps": rate,
        "queries": n_queries,
        "achieved_qps": n_queries / elapsed,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "submit_span_s": t_submitted,
        "batches": st["batches"],
        "mean_batch_queries": st["mean_batch_queries"],
        "deferred": st["deferred"],
        "reruns": st["reruns"],
        "peak_held_gb": st["peak_held_bytes"] / 1e9,
        "host_plan_ms_per_batch": st["host_plan_ms"] / max(st["batches"], 1),
        "host_finish_ms_per_batch": st["host_finish_ms"] / max(st["batches"], 1),
        "resident_gb": st["resident_bytes"] / 1e9,
        "server_device_gb": st["device_bytes"] / 1e9,
    }


def bench_serving(args):
    """Fixed-QPS sweeps of the batching server and of the per-query baseline;
    each side's highest offered rate whose p99 stays within the target."""
    import torch

    log(f"generating {args.tables} serving tables")
    tables, rows, card = make_c5_tables(args.tables, args.seed, max_rows=args.max_table_rows)
    torch.cuda.synchronize()
    budget = int(args.hbm_budget_gb * (1 << 30))
    out = {"tables": args.tables, "table_rows_total": int(rows.sum()), "hbm_budget_gb": args.hbm_budget_gb,
           "p99_target_ms": args.p99_target_ms, "runs": {"batched": [], "per_query": []}}
    for mode, rates in (("per_query", args.base_rates), ("batched", args.batch_rates)):
        for rate in rates:
            n = int(min(args.serve_max_queries, max(args.serve_min_queries, rate * args.serve_seconds)))
            log(f"serving {mode}: {rate} q/s, {n} queries")
            r = serve_run(tables, rate, n, mode == "batched", args.seed + int(rate), budget,
                          reserve=int(args.server_reserve_gb * (1 << 30)))
            log(f"  p50 {r['p50_ms']:.2f} ms  p99 {r['p99_ms']:.2f} ms  achieved {r['achieved_qps']:.0f} q/s  "
                f"batches {r['batches']} (mean {r['mean_batch_queries']:.1f}); host plan "
                f"{r['host_plan_ms_per_batch']:.2f} ms, finish {r['host_finish_ms_per_batch']:.2f} ms per batch")
            out["runs"][mode].append(r)
            if r["p99_ms"] > args.p99_target_ms or r["achieved_qps"] < 0.8 * rate:
                break
    for mode in ("batched", "per_query"):
        ok = [r for r in out["runs"][mode] if r["p99_ms"] <= args.p99_target_ms and r["achieved_qps"] >= 0.8 * r["offered_qps"]]
        out[f"{mode}_qps_at_target"] = max((r["achieved_qps"] for r in ok), default=0.0)
        out[f"{mode}_p99_at_target"] = min((r["p99_ms"] for r in ok if r["achieved_qps"] == out[f"{mode}_qps_at_target"]),
                                           default=None)
    b, p = out["batched_qps_at_target"], out["per_query_qps_at_target"]
    out["speedup_at_equal_p99"] = b / p if p else None
    return out


def bench_pandas(args):
    """pandas' df.groupby("key").agg(spec) through the hook (cudf_amd.pandas_accel,
    the cudf.pandas-style path) on a host DataFrame of the headline distribution,
    end to end: columns copied to HBM, the groupby, results copied back and
    assembled as pandas' result. pandas' own call is timed beside it."""
    import numpy as np
    import pandas as pd

    from cudf_amd import pandas_accel as PA

    rows = args.rows or 10_000_000
    rng = np.random.default_rng(args.seed)
    uniq = np.unique(rng.integers(-2 ** 63, 2 ** 63 - 1, args.distinct, dtype=np.int64))
    df = pd.DataFrame({"key": uniq[rng.integers(0, uniq.size, rows)],
                       "f64": rng.standard_normal(rows), "f32": rng.standard_normal(rows).astype(np.float32),
                       "i32": rng.integers(-1_000_000, 1_000_001, rows).astype(np.int32)})
    for c in ("f64", "f32"):
        df.loc[rng.random(rows) < 0.05, c] = np.nan
    names = args.pandas_aggs.split(",")
    spec = {"f64": names, "f32": names, "i32": names}
    for _ in range(max(args.warmup, 1)):
        PA.groupby_agg(df, "key", spec)
    t_gpu = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        PA.groupby_agg(df, "key", spec)
        t_gpu.append(time.perf_counter() - t0)
    t_pd = []
    for _ in range(3):
        t0 = time.perf_counter()
        df.groupby("key").agg(spec)
        t_pd.append(time.perf_counter() - t0)
    g, p = statistics.median(t_gpu), statistics.median(t_pd)
    return {"rows": rows, "hook_s": g, "pandas_s": p, "hook_rows_per_s": rows / g, "pandas_rows_per_s": rows / p,
            "speedup": p / g}


def bench_batched(args):
    """4096 independent queries: one planned batch (one HIP graph per
    execution) against the same queries as sequential plain calls."""
    import numpy as np
    import torch

    import cudf_amd as ca

    Q = args.queries
    log(f"generating {Q} batched queries")
    queries, rows, card = make_c4(Q, args.seed, max_rows=args.max_query_rows)
    torch.cuda.synchronize()
    total_rows = int(rows.sum())
    log(f"{total_rows:,} rows in total; planning the batch")
    t0 = time.perf_counter()
    batch = ca.GroupByBatch(queries)
    plan_s = time.perf_counter() - t0
    stats = batch.stats()
    log(f"plan {plan_s * 1e3:.1f} ms: {stats}")
    for _ in range(args.warmup):
        batch.run()
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run()   # execute (graph launch) + finish (wait, group counts)
        times.append(time.perf_counter() - t0)
    t_batch = statistics.median(times)
    groups = [batch.num_groups(q) for q in range(Q)]
    log(f"batched: median {t_batch * 1e3:.2f} ms per batch of {Q}")

    # the same queries one plain groupby call at a time
    n_base = min(Q, args.baseline_queries)
    log(f"per-query baseline over {n_base} queries")
    for q in range(min(8, n_base)):
        ca.GroupBy(queries[q][0]).aggregate(queries[q][1])
    torch.cuda.synchronize()
    lat = []
    t_all = time.perf_counter()
    for q in range(n_base):
        t0 = time.perf_counter()
        k, r = ca.GroupBy(queries[q][0]).aggregate(queries[q][1])
        torch.cuda.synchronize()
        lat.append(time.perf_counter() - t0)
        del k, r
    t_base = time.perf_counter() - t_all
    base_qps = n_base / t_base
    in_bytes = 12.0 * total_rows
    return {
        "queries": Q,
        "rows_total": total_rows,
        "batch_ms": t_batch * 1e3,
        "batch_ms_all": [round(t * 1e3, 3) for t in times],
        "plan_ms": plan_s * 1e3,
        "qps": Q / t_batch,
        "rows_per_s": total_rows / t_batch,
        "hbm_frac": in_bytes / t_batch / 1e9 / HBM_PEAK_GBS,
        "groups_total": int(sum(groups)),
        "modes": {"lds_direct": stats["direct_queries"], "global": stats["global_queries"],
                  "partitioned": stats["partitioned_queries"]},
        "reruns": batch.stats()["reruns"],
        "per_query_baseline": {
            "queries": n_base,
            "qps": base_qps,
            "p50_ms": float(np.percentile(lat, 50) * 1e3),
            "p99_ms": float(np.percentile(lat, 99) * 1e3),
        },
        "speedup_vs_per_query": (Q / t_batch) / base_qps,
        "latency_note": "every query of a batch completes with the batch: its latency is batch_ms",
    }


def oracle_sample(name: str, sample_rows: int, distinct: int, seed: int):
    """Host columns of a bounded sample drawn like the workload, for the
    oracle: (keys, requests)."""
    import numpy as np

    from oracle import oracle as O

    rng = np.random.default_rng(seed)
    n = sample_rows
    kinds5 = [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN]
    if name == "c2v":
        kinds5 = [O.MEAN, (O.VARIANCE, 1), (O.STD, 1), O.ARGMIN, O.ARGMAX]
    if name == "c2s":
        kinds5 = [(O.MEDIAN, 0, 0, 0, 0.5), (O.QUANTILE, 0, 0, 0, 0.9), (O.NUNIQUE, 0, 0, 0, 0.5),
                  (O.NTH_ELEMENT, 0, 1, 0, 0.5)]
    if name in ("c2", "c2v", "c2s"):
        uniq = np.unique(rng.integers(-2 ** 63, 2 ** 63 - 1, distinct, dtype=np.int64))
        keys = [O.HostColumn(O.INT64, uniq[rng.integers(0, uniq.size, n)])]
        reqs = [(O.HostColumn(O.INT32, rng.integers(-1_000_000, 1_000_001, n).astype(np.int32), rng.random(n) >= 0.05),
                 kinds5),
                (O.HostColumn(O.FLOAT64, rng.standard_normal(n), rng.random(n) >= 0.05), kinds5),
                (O.HostColumn(O.FLOAT32, rng.standard_normal(n).astype(np.float32), rng.random(n) >= 0.05), kinds5)]
    elif name == "c1":
        keys = [O.HostColumn(O.INT32, rng.integers(0, 1000, n).astype(np.int32))]
        reqs = [(O.HostColumn(O.FLOAT64, rng.random(n)), kinds5)]
    elif name == "c3a":
        n_keys = 50_000_000
        u = rng.random(n)
        r = np.clip(np.floor((1.0 + u * (n_keys ** -0.1 - 1.0)) ** -10.0).astype(np.int64) - 1, 0, n_keys - 1)
        keys = [O.HostColumn(O.INT32, (r * 2654435761 % 2 ** 31).astype(np.int32)),
                O.HostColumn(O.INT64, (r.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
                                       % np.uint64(2 ** 61)).astype(np.int64))]
        reqs = [(O.HostColumn(O.FLOAT64, rng.standard_normal(n)), [O.SUM, O.COUNT_VALID, O.MAX])]
    else:  # c3b
        uniq = rng.integers(-2 ** 31, 2 ** 31 - 1, 16).astype(np.int32)
        keys = [O.HostColumn(O.INT32, uniq[rng.integers(0, 16, n)])]
        reqs = [(O.HostColumn(O.FLOAT64, rng.standard_normal(n)), [O.SUM, O.COUNT_VALID, O.MAX])]
    return keys, reqs


def cpu_baseline(name: str, pandas_rows: int, oracle_rows: int, distinct: int, seed: int):
    """The reported CPU baseline: pandas' groupby-agg, which is what the
    reference runs on the CPU (the cudf.pandas fallback executes pandas,
    python/cudf/cudf/pandas/module_accelerator.py:79-92), on a bounded sample
    drawn like the workload: default sort, 2 warm-ups, median of 7 runs, one
    core. Beside it, the oracle (the C restatement of the reference's hash
    groupby, oracle/groupby_oracle.c, one thread) on its own sample."""
    from oracle import oracle as O

    if name in ("c2", "c2v", "c2s"):
        aggs = {"c2": ("sum", "count", "min", "max", "mean"), "c2v": ("mean", "var", "std", "idxmin", "idxmax"),
                "c2s": ("median", "nunique", "first")}[name]
        pd_line = cpu_baseline_c2(pandas_rows, distinct, seed, aggs=aggs)
    else:
        pd_line = cpu_baseline_generic(name, pandas_rows, seed)
    O.load()
    keys, reqs = oracle_sample(name, oracle_rows, distinct, seed)
    t0 = time.perf_counter()
    O.groupby(keys, reqs)
    t = time.perf_counter() - t0
    pd_line["oracle_port"] = {
        "value": oracle_rows / t, "unit": "rows/s", "cores": 1, "kind": "port",
        "sample": f"the oracle's C groupby (oracle/groupby_oracle.c, one thread) on {oracle_rows:,} rows drawn like "
                  f"the workload, one call of {t:.1f} s"}
    return pd_line


def _time_pandas(run, warmups: int = 2, reps: int = 7) -> float:
    for _ in range(warmups):
        run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    return statistics.median(times)


def cpu_baseline_generic(name: str, sample_rows: int, seed: int):
    """pandas groupby-agg on a bounded sample drawn like the workload."""
    import numpy as np
    import pandas as pd

    rng = np.random.default_rng(seed)
    if name == "c1":
        df = pd.DataFrame({"key": rng.integers(0, 1000, sample_rows).astype(np.int32), "val": rng.random(sample_rows)})
        run = lambda: df.groupby("key")["val"].agg(["sum", "count", "mean", "min", "max"])
        what = "int32 key in [0, 1000), float64 U[0, 1); sum/count/mean/min/max"
    elif name == "c3a":
        n_keys = 50_000_000
        u = rng.random(sample_rows)
        r = np.clip(np.floor((1.0 + u * (n_keys ** -0.1 - 1.0)) ** -10.0).astype(np.int64) - 1, 0, n_keys - 1)
        df = pd.DataFrame({"a": (r * 2654435761 % 2 ** 31).astype(np.int32),
                           "b": (r.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15) % np.uint64(2 ** 61)).astype(np.int64),
                           "val": rng.standard_normal(sample_rows)})
        run = lambda: df.groupby(["a", "b"])["val"].agg(["sum", "count", "max"])
        what = "(int32, int64) Zipf s=1.1 keys, float64 value; sum/count/max"
    else:  # c3b
        uniq = rng.integers(-2 ** 31, 2 ** 31 - 1, 16).astype(np.int32)
        df = pd.DataFrame({"key": uniq[rng.integers(0, 16, sample_rows)], "val": rng.standard_normal(sample_rows)})
        run = lambda: df.groupby("key")["val"].agg(["sum", "count", "max"])
        what = "int32 key over 16 values, float64 value; sum/count/max"
    t = _time_pandas(run)
    return {"value": sample_rows / t, "unit": "rows/s", "cores": 1, "kind": "reference",
            "sample": f"pandas {pd.__version__} groupby-agg (the reference's CPU path: cudf.pandas falls back to "
                      f"pandas) on {sample_rows:,} rows drawn like the workload ({what}), default sort, 2 warm-ups, "
                      f"median of 7, single-threaded; host os.cpu_count()={os.cpu_count()}"}


def cpu_baseline_c2(sample_rows: int, distinct: int, seed: int, aggs=("sum", "count", "min", "max", "mean")):
    """pandas groupby-agg on a bounded sample of the 100M-row workload's
    distribution (nullable columns; int32 values summed as Int64, as cudf
    widens them)."""
    import numpy as np
    import pandas as pd

    rng = np.random.default_rng(seed)
    uniq = np.unique(rng.integers(-2 ** 63, 2 ** 63 - 1, distinct, dtype=np.int64))
    keys = uniq[rng.integers(0, uniq.size, sample_rows)]
    df = pd.DataFrame({
        "key": keys,
        "i32": pd.array(rng.integers(-1_000_000, 1_000_001, sample_rows).astype(np.int32), dtype="Int64"),
        "f64": pd.array(rng.standard_normal(sample_rows), dtype="Float64"),
        "f32": pd.array(rng.standard_normal(sample_rows).astype(np.float32), dtype="Float32"),
    })
    for c in ("i32", "f64", "f32"):
        df.loc[rng.random(sample_rows) < 0.05, c] = pd.NA
    aggs = list(aggs)
    t = _time_pandas(lambda: df.groupby("key").agg({"i32": aggs, "f64": aggs, "f32": aggs}))
    return {
        "value": sample_rows / t,
        "unit": "rows/s",
        "cores": 1,
        "kind": "reference",
        "sample": (f"pandas {pd.__version__} groupby-agg (the reference's CPU path: cudf.pandas falls back to pandas) "
                   f"on {sample_rows:,} rows of the same distribution ({distinct:,} distinct int64 keys, 3 value "
                   f"columns 5% null, {'/'.join(aggs)} each), default sort, 2 warm-ups, median of 7, "
                   f"single-threaded; host os.cpu_count()={os.cpu_count()}"),
    }


def summarize_trace(db_path: str, title: str, command: str) -> str:
    """Text summary of the top_kernels view of a rocprofv3 --kernel-trace --stats database."""
    import sqlite3

    con = sqlite3.connect(db_path)
    rows = list(con.execute("select name, total_calls, total_duration, average, percentage from top_kernels"))
    out = [title, "Command (from the repo root, TMPDIR=/tmp):", "  " + command,
           "Durations in microseconds (top_kernels view of the rocprofv3 database; torch kernels are data generation).",
           "", f"{'kernel':60s} {'calls':>7s} {'total_us':>12s} {'avg_us':>10s} {'pct':>6s}"]
    for name, calls, total, avg, pct in rows[:30]:
        out.append(f"{kernel_name(name)[:60]:60s} {int(calls):7d} {float(total):12.1f} {float(avg):10.1f} "
                   f"{float(pct):6.2f}")
    return "\n".join(out) + "\n"


def summarize_pmc(fetch_csv: str, write_csv: str, cal_fetch_csv: str, cal_write_csv: str, title: str,
                  commands) -> str:
    """Per-kernel HBM bytes per dispatch from FETCH_SIZE / WRITE_SIZE passes,
    each scaled by what the counter reads for a copy of 1 GiB (read and
    written once)."""
    import collections
    import csv

    def per_kernel(path, counter):
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                acc[kernel_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        return acc

    cal_f = per_kernel(cal_fetch_csv, "FETCH_SIZE")
    cal_w = per_kernel(cal_write_csv, "WRITE_SIZE")
    copy_name = next(k for k in cal_f if "copyBuffer" in k or "copy" in k.lower())
    gib_kib = float(1 << 20)
    f_scale = gib_kib / statistics.median(cal_f[copy_name])
    w_scale = gib_kib / statistics.median(cal_w[copy_name])
    fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    write = per_kernel(write_csv, "WRITE_SIZE")
    out = [title, "Commands (from the repo root, TMPDIR=/tmp):"] + ["  " + c for c in commands] + [
        f"Calibration: a 1 GiB device copy ({copy_name}) reads FETCH_SIZE = {statistics.median(cal_f[copy_name]):.0f} KiB "
        f"and WRITE_SIZE = {statistics.median(cal_w[copy_name]):.0f} KiB per dispatch: FETCH_SIZE is scaled by "
        f"{f_scale:.3f}, WRITE_SIZE by {w_scale:.3f}.", "",
        f"{'kernel':48s} {'dispatches':>10s} {'read_MB':>10s} {'written_MB':>11s}"]
    for k in sorted(fetch, key=lambda k: -statistics.mean(fetch[k])):
        if not k.startswith("mgb::"):
            continue
        rd = statistics.mean(fetch[k]) * f_scale * 1024 / 1e6
        wr = statistics.mean(write.get(k, [0.0])) * w_scale * 1024 / 1e6
        out.append(f"{k[:48]:48s} {len(fetch[k]):10d} {rd:10.1f} {wr:11.1f}")
    return "\n".join(out) + "\n"


def summarize_counters(csv_paths, title: str, commands) -> str:
    """Per-kernel mean of every counter of one or more rocprofv3 --pmc CSV
    files (one per counter pass), for the library's kernels."""
    import collections
    import csv

    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for path in csv_paths:
        for r in csv.DictReader(open(path)):
            name = kernel_name(r["Kernel_Name"])
            if name.startswith("mgb::"):
                acc[name][r["Counter_Name"]].append(float(r["Counter_Value"]))
    counters = sorted({c for k in acc.values() for c in k})
    out = [title, "Commands (from the repo root, TMPDIR=/tmp):"] + ["  " + c for c in commands] + [
        "Mean value per dispatch of each counter (summed over the device's shader engines by rocprofv3).", ""]
    for k in sorted(acc):
        out.append(f"{k}  (dispatches: {max(len(v) for v in acc[k].values())})")
        for c in counters:
            if c in acc[k]:
                out.append(f"    {c:28s} {statistics.mean(acc[k][c]):18.0f}")
    return "\n".join(out) + "\n"


def find_kernel(path: str, scope: str, default: str) -> str:
    """The kernel that a profiling scope launched: the compact execution's
    instance named in the traffic summary when there is one, else `default`."""
    prefix = COMPACT_PREFIX.get(scope)
    try:
        for line in open(path):
            parts = line.split()
            if prefix and parts and parts[0].startswith(prefix):
                return parts[0]
    except OSError:
        pass
    return default


def read_traffic(path: str, kernel: str):
    """HBM bytes per dispatch of `kernel` from a summary written by summarize_pmc, or None."""
    try:
        for line in open(path):
            parts = line.split()
            if parts and parts[0] == kernel and len(parts) >= 4:
                return (float(parts[2]) + float(parts[3])) * 1e6
    except OSError:
        return None
    return None


def read_call_traffic(path: str, kernel: str):
    """HBM bytes of one whole call from a summary written by summarize_pmc:
    every kernel's bytes per dispatch times its dispatches per call (the
    dispatches of `kernel`, which runs once per call, count the calls)."""
    try:
        rows = [line.split() for line in open(path)]
    except OSError:
        return None
    rows = [r for r in rows if len(r) == 4 and r[0].startswith("mgb::") and r[1].isdigit()]
    calls = next((int(r[1]) for r in rows if r[0] == kernel), 0)
    if not calls:
        return None
    return sum((float(r[2]) + float(r[3])) * 1e6 * int(r[1]) for r in rows) / calls


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c2v", "c2s", "c3a", "c3b", "c4", "c5", "pandas"])
    ap.add_argument("--tables", type=int, default=32, help="c5: resident tables")
    ap.add_argument("--max-table-rows", type=float, default=5e7, help="c5: largest table (rows)")
    ap.add_argument("--hbm-budget-gb", type=float, default=200.0, help="c5: HBM budget of the server (GiB)")
    ap.add_argument("--p99-target-ms", type=float, default=100.0, help="c5: p99 latency target")
    ap.add_argument("--server-reserve-gb", type=float, default=64.0,
                    help="c5: device memory the server takes at creation for batches and results (GiB)")
    ap.add_argument("--base-rates", type=float, nargs="*", default=[400, 800, 1200, 1600, 2000, 2400])
    ap.add_argument("--batch-rates", type=float, nargs="*", default=[800, 1600, 2400, 3200, 4000, 4800, 5600, 6400])
    ap.add_argument("--serve-seconds", type=float, default=2.0)
    ap.add_argument("--serve-min-queries", type=int, default=200)
    ap.add_argument("--serve-max-queries", type=int, default=4000)
    ap.add_argument("--queries", type=int, default=4096, help="c4: queries per batch")
    ap.add_argument("--max-query-rows", type=float, default=1e7, help="c4: largest query (rows)")
    ap.add_argument("--baseline-queries", type=int, default=4096, help="c4: queries timed one plain call each")
    ap.add_argument("--rows", type=int, default=None)
    ap.add_argument("--distinct", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-sample-rows", type=int, default=10_000_000,
                    help="rows of the oracle's timed CPU sample (about 2 s for the 100M-row workload)")
    ap.add_argument("--pandas-sample-rows", type=int, default=10_000_000,
                    help="rows of the reference CPU path's (pandas) timed sample (9 runs, about 12 s at 10M rows)")
    ap.add_argument("--settle-s", type=float, default=1.0,
                    help="seconds of untimed calls before the warm-up steps (a fresh process's first calls "
                         "include one-time host stalls)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pandas-aggs", default="sum,count,min,max,mean",
                    help="pandas workload: the aggregations of each value column (pandas names)")
    ap.add_argument("--summarize-trace", nargs=3, metavar=("RESULTS_DB", "TITLE", "COMMAND"),
                    help="print a text summary of a rocprofv3 --kernel-trace --stats database and exit")
    ap.add_argument("--summarize-pmc", nargs=6,
                    metavar=("FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV", "CAL_WRITE_CSV", "TITLE", "COMMANDS"),
                    help="print per-kernel HBM bytes from FETCH_SIZE / WRITE_SIZE passes and exit "
                         "(COMMANDS: the profiling commands, separated by ';;')")
    ap.add_argument("--summarize-counters", nargs=3, metavar=("CSV_LIST", "TITLE", "COMMANDS"),
                    help="print per-kernel means of the counters in rocprofv3 --pmc CSV files and exit "
                         "(CSV_LIST and COMMANDS separated by ';;')")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "hbm_traffic_groupby_c2_100M_rows.txt"))
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only copy a 1 GiB buffer 3 times (reference kernel for FETCH_SIZE / WRITE_SIZE scaling)")
    args = ap.parse_args()
    if args.summarize_trace:
        print(summarize_trace(*args.summarize_trace), end="")
        return
    if args.summarize_pmc:
        a = args.summarize_pmc
        print(summarize_pmc(a[0], a[1], a[2], a[3], a[4], a[5].split(";;")), end="")
        return
    if args.summarize_counters:
        a = args.summarize_counters
        print(summarize_counters(a[0].split(";;"), a[1], a[2].split(";;")), end="")
        return
    if args.calibrate_copy:
        import torch

        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": 1 << 30, "bytes_written": 1 << 30}))
        return

    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU and "
              "this groupby does not shard across GPUs in this build)", file=sys.stderr)
        sys.exit(2)

    import torch

    import cudf_amd as ca
    from cudf_amd import _lib
    import ctypes

    torch.cuda.set_device(0)
    lib = _lib.load()
    if args.workload == "pandas":
        res = bench_pandas(args)
        print(json.dumps({
            "metric": METRIC, "value": res["hook_rows_per_s"], "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["hook_s"] * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int64 key; f64/f32/int32 values; f64 and int64 accumulation",
            "data": "synthetic (seeded host DataFrame)",
            "config": {"workload": f"pandas df.groupby('key').agg(...) through the hook, {res['rows']:,} rows, "
                                   f"{args.distinct:,} distinct int64 keys, f64/f32 values 5% NaN, int32 values, "
                                   f"{'/'.join(args.pandas_aggs.split(','))} each; host DataFrame in and out "
                                   "(PCIe inclusive)",
                       "parallelism": "single GPU"},
            "cpu_baseline": {"value": res["pandas_rows_per_s"], "unit": "rows/s", "cores": 1, "kind": "port",
                             "sample": f"pandas groupby-agg of the same DataFrame, median of 3"},
            "pandas": res}), flush=True)
        return
    if args.workload == "c4":
        res = bench_batched(args)
        print(json.dumps({
            "metric": METRIC, "value": res["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["batch_ms"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32 key; f64 value; f64 accumulation", "data": "synthetic (seeded on device)",
            "config": {"workload": f"batched groupby: {args.queries} queries, rows log-uniform [1e3,"
                                   f"{args.max_query_rows:.0e}], groups log-uniform [10,1e5], int32 key, f64 value, "
                                   "sum/count/mean", "parallelism": "single GPU"},
            "batched": res}), flush=True)
        return
    if args.workload == "c5":
        res = bench_serving(args)
        print(json.dumps({
            "metric": METRIC, "value": res["batched_qps_at_target"], "unit": "queries/s at p99 target",
            "n_gpus": 1, "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "int64 key; f64 + int32 values; f64/int64 accumulation",
            "data": "synthetic (seeded on device)",
            "config": {"workload": f"serving: Poisson arrivals over {args.tables} resident tables, rows log-uniform "
                                   f"[1e4,{args.max_table_rows:.0e}], groups log-uniform [10,1e6], int64 key, f64 + "
                                   "int32 values 2% null, sum/count/min/max/mean, HBM budget "
                                   f"{args.hbm_budget_gb:.0f} GiB", "parallelism": "single GPU"},
            "serving": res}), flush=True)
        return
    rows = args.rows or DEFAULT_ROWS[args.workload]
    args.rows = rows
    log(f"generating workload {args.workload} ({rows:,} rows)")
    key_cols, reqs, in_bytes_per_row, out_bytes_per_group, description = make_workload(args.workload, rows, args.seed, distinct=args.distinct)
    torch.cuda.synchronize()
    gb = ca.GroupBy(ca.Table(key_cols))
    groups = 0
    log(f"settling ({args.settle_s:.1f} s of untimed calls)")
    t_end = time.perf_counter() + args.settle_s
    while True:
        k, r = gb.aggregate(reqs)
        groups = k.num_rows()
        del k, r
        torch.cuda.synchronize()
        if time.perf_counter() >= t_end:
            break
    log("warmup")
    for _ in range(args.warmup):
        k, r = gb.aggregate(reqs)
        groups = k.num_rows()
        del k, r
    torch.cuda.synchronize()

    log(f"timing {args.steps} steps")
    stream = torch.cuda.current_stream()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record(stream)
    for _ in range(args.steps):
        k, r = gb.aggregate(reqs)
        groups = k.num_rows()
        del k, r
    stop.record(stream)
    torch.cuda.synchronize()
    total_ms = start.elapsed_time(stop)

    # per-kernel durations from a separate pass of the same steps, with HIP
    # events around every launch on its stream (not in the timed region)
    log("kernel timings")
    lib.mgb_profiling_enable(1)
    for name in PROFILED:
        n, ms = ctypes.c_int64(), ctypes.c_double()
        lib.mgb_profiling_read(name.encode(), ctypes.byref(n), ctypes.byref(ms))  # drop stale records
    for _ in range(args.steps):
        k, r = gb.aggregate(reqs)
        del k, r
    torch.cuda.synchronize()
    kern = {}
    for name in PROFILED:
        n, ms = ctypes.c_int64(), ctypes.c_double()
        _lib.check(lib.mgb_profiling_read(name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        if n.value:
            kern[name] = (n.value, ms.value)
    lib.mgb_profiling_enable(0)

    R = args.rows
    ms_per_step = total_ms / args.steps
    # c2 per group: key 8 + int32 col (sum 8, count 4, min 4, max 4, mean 8) + f64 col (8, 4, 8, 8, 8)
    # + f32 col (4, 4, 4, 4, 8) = 96 bytes, plus 12 validity bits
    in_bytes = R * in_bytes_per_row
    out_bytes = groups * out_bytes_per_group
    # the dominant kernel: the longest of the passes that each read every row
    # once (its algorithmic bytes: the input columns, read once)
    dom = max((k for k in ("aggregate", "lds_aggregate", "partition_scatter", "pass2", "sort_aggs") if k in kern),
              key=lambda k: kern[k][1] / kern[k][0])
    agg_n, agg_ms = kern[dom]
    symbol = find_kernel(args.traffic_file, dom, KERNEL_SYMBOL[dom]) if args.workload == "c2" else KERNEL_SYMBOL[dom]
    if dom == "lds_aggregate" and args.workload == "c2v":
        symbol = "mgb::lds_agg_kernel<true>"  # the instance with the second pass
    agg_avg_ms = agg_ms / agg_n
    achieved = in_bytes / (agg_avg_ms * 1e-3) / 1e9
    traffic = read_traffic(args.traffic_file, symbol) if args.workload == "c2" else None
    call_gbs = (in_bytes + out_bytes) / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline")
        cpu = cpu_baseline(args.workload, min(args.pandas_sample_rows, R), min(args.cpu_sample_rows, R),
                           args.distinct, args.seed)

    line = {
        "metric": METRIC,
        "value": R / (ms_per_step * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int key(s); int/float values; int64 and f64 accumulation",
        "data": "synthetic (seeded on device)",
        "config": {"workload": description, "rows": R, "groups": groups, "parallelism": "single GPU"},
        # the whole call: the algorithmic bytes (input columns read once,
        # results written once) over the measured time per step
        "roofline": {
            "bound": "hbm",
            "scope": "whole call (inputs read once + results written once, per step)",
            "achieved": call_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": call_gbs / HBM_PEAK_GBS,
            "traffic": read_call_traffic(args.traffic_file, symbol) if args.workload == "c2" else None,
            "algorithmic_bytes_per_step": in_bytes + out_bytes,
            # the dominant kernel: its algorithmic bytes per launch (the input
            # columns, read once) over its average launch duration
            "dominant_kernel": {
                "kernel": symbol,
                "achieved": achieved,
                "frac": achieved / HBM_PEAK_GBS,
                "traffic": traffic,
                "traffic_source": os.path.relpath(args.traffic_file, ROOT) if traffic is not None else None,
                "algorithmic_bytes_per_launch": in_bytes,
                "avg_launch_ms": agg_avg_ms,
            },
        },
        "kernel_ms": {k: (v[1] / max(v[0], 1)) for k, v in kern.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
