"""Benchmark of the MI355X hash groupby-aggregate.

Default workload (the headline): 100M rows, int64 key drawn from 1M distinct
values spread over the whole int64 range; value columns int32 uniform in
[-1e6, 1e6], float64 N(0,1) and float32 N(0,1), each 5% null; sum, count,
min, max and mean of every value column. One step = one
cudf::groupby::groupby::aggregate call through the C ABI on device-resident
inputs (output allocation included).

Prints one JSON line. `value` is rows/s over the timed steps; `roofline`
prices the dominant kernel (aggregate) with its algorithmic bytes (the input
columns read once) over its own average duration, measured with HIP events on
its launch stream; `cpu_baseline` times pandas (the cudf.pandas CPU fallback)
on a bounded sample of the same data on this host.
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
HBM_PEAK_GBS = 8000.0


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


def cpu_baseline_c2(sample_rows: int, distinct: int, seed: int, reps: int = 3):
    """pandas groupby-agg on a bounded sample of the config-2 distribution."""
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
    aggs = ["sum", "count", "min", "max", "mean"]
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        df.groupby("key", sort=False).agg({"i32": aggs, "f64": aggs, "f32": aggs})
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {
        "value": sample_rows / t,
        "unit": "rows/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"pandas {pd.__version__} groupby-agg (the cudf.pandas CPU fallback) on {sample_rows:,} rows "
                   f"of the same distribution ({distinct:,} distinct int64 keys, 3 value columns 5% null, "
                   f"5 aggs each), median of {reps}, single-threaded; host os.cpu_count()={os.cpu_count()}"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=100_000_000)
    ap.add_argument("--distinct", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--cpu-sample-rows", type=int, default=10_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

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
    key_col, cols, distinct = make_c2(args.rows, args.distinct, args.seed)
    torch.cuda.synchronize()
    reqs = c2_requests(cols)
    gb = ca.GroupBy(ca.Table([key_col]))
    groups = 0
    for _ in range(args.warmup):
        k, r = gb.aggregate(reqs)
        groups = k.num_rows()
        del k, r
    torch.cuda.synchronize()

    lib.mgb_profiling_enable(1)
    for name in (b"aggregate", b"fill", b"finalize", b"partition_hist", b"partition_scatter"):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        lib.mgb_profiling_read(name, ctypes.byref(n), ctypes.byref(ms))  # drop stale records
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
    kern = {}
    for name in ("aggregate", "fill", "finalize", "partition_hist", "partition_scatter"):
        n, ms = ctypes.c_int64(), ctypes.c_double()
        _lib.check(lib.mgb_profiling_read(name.encode(), ctypes.byref(n), ctypes.byref(ms)))
        kern[name] = (n.value, ms.value)
    lib.mgb_profiling_enable(0)

    R = args.rows
    ms_per_step = total_ms / args.steps
    in_bytes_per_row = 8 + 4 + 8 + 4 + 3 / 8
    in_bytes = R * in_bytes_per_row
    # per group: key 8 + int32 col (sum 8, count 4, min 4, max 4, mean 8) + f64 col (8, 4, 8, 8, 8)
    # + f32 col (4, 4, 4, 4, 8) = 96 bytes, plus 12 validity bits
    out_bytes = groups * 96 + groups * 12 / 8
    agg_n, agg_ms = kern["aggregate"]
    agg_avg_ms = agg_ms / max(agg_n, 1)
    achieved = in_bytes / (agg_avg_ms * 1e-3) / 1e9
    call_gbs = (in_bytes + out_bytes) / (ms_per_step * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_c2(min(args.cpu_sample_rows, R), args.distinct, args.seed)

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
        "dtype": "int64 key; int32/f64/f32 values (f64 accumulation)",
        "data": "synthetic (seeded on device)",
        "config": {
            "workload": "groupby-aggregate, 100M rows, int64 key over 1M distinct values, int32/f64/f32 values "
                        "5% null, sum/count/min/max/mean per value column",
            "rows": R, "distinct_keys": distinct, "groups": groups, "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "aggregate_kernel",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": in_bytes,
            "avg_launch_ms": agg_avg_ms,
        },
        "whole_call": {
            "algorithmic_bytes": in_bytes + out_bytes,
            "achieved_GBs": call_gbs,
            "frac": call_gbs / HBM_PEAK_GBS,
            "kernel_ms": {k: (v[1] / max(v[0], 1)) for k, v in kern.items()},
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
