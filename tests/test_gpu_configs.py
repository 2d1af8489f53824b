"""GPU parity at the benchmark workloads' own shapes:

* the CPU workload (1M rows, int32 key in [0, 1000), float64 U[0, 1), sum /
  count / mean / min / max) against the oracle and against pandas, the
  reference's own CPU path (the cudf.pandas fallback runs pandas);
* the 100M-row headline workload (int64 key over 1M distinct values, int32 /
  float64 / float32 values each 5% null, five aggregations each) against the
  oracle at full size;
* a 4096-query batch drawn like the batched workload (rows log-uniform in
  [1e3, 1e7], groups log-uniform in [10, 1e5], 4.4e9 rows in all): every
  query equal to its own plain call, its group count equal to its distinct
  keys, a sample against the oracle;
* a served mix over tables of up to 3e7 rows and 1e6 groups: every query's
  result equal to the same query run alone, one large table against the
  oracle.

Integers are bit-exact; float64 within 1e-12 relative, float32 within 1e-5
(or the same scaled by the group's sum of |x|)."""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest
import torch

import bench
import cudf_amd as ca
from cudf_amd import aggregation as agg
from oracle import oracle as O
from tests.helpers import assert_same, gpu_groupby, oracle_groupby, sorted_result

pytestmark = pytest.mark.gpu

KINDS = [O.SUM, O.COUNT_VALID, O.MEAN, O.MIN, O.MAX]


def progress(msg):
    print(f"[configs] {msg}", file=sys.__stderr__, flush=True)


def _compare_to_oracle(keys, reqs, rtols):
    """keys / reqs as tests.helpers takes them; the abs-sum tolerance comes
    from the same oracle call (SUM of |x| per request column)."""
    gk, gr = gpu_groupby(keys, reqs)
    progress("gpu done; oracle")
    abs_reqs = [((O.FLOAT64, np.abs(np.asarray(v, np.float64)), valid), [O.SUM]) for (t, v, valid), _ in reqs]
    ok, orr = oracle_groupby(keys, list(reqs) + abs_reqs)
    progress("oracle done; comparing")
    gk, gr = sorted_result(gk, gr)
    ok, orr = sorted_result(ok, orr)
    assert len(gk[0][0]) == len(ok[0][0]), f"{len(gk[0][0])} groups vs {len(ok[0][0])}"
    assert np.array_equal(gk[0][0], ok[0][0]), "group keys differ"
    n = len(reqs)
    for r in range(n):
        s = orr[n + r][0][0]
        kinds = list(reqs[r][1])
        count = orr[r][kinds.index(O.COUNT_VALID)][0] if O.COUNT_VALID in kinds else None
        for a, (g, o) in enumerate(zip(gr[r], orr[r])):
            # SUM scaled by the group's sum of |x|, MEAN by its mean of |x|
            scale = s if kinds[a] == O.SUM else \
                s / np.maximum(count, 1) if kinds[a] == O.MEAN and count is not None else None
            assert_same(g, o, None, abs_sum=scale, rtol=rtols[r], what=f"request {r} agg {kinds[a]}")
    return gk, gr


def test_c1_cpu_workload_exact(gpu):
    """1M rows, int32 key uniform in [0, 1000), float64 uniform [0, 1), no
    nulls: sum, count, mean, min, max; against the oracle and pandas."""
    import pandas as pd

    rng = np.random.default_rng(42)
    n = 1_000_000
    key = rng.integers(0, 1000, n).astype(np.int32)
    val = rng.random(n)
    gk, gr = _compare_to_oracle([(O.INT32, key, None)], [((O.FLOAT64, val, None), KINDS)], [None])
    p = pd.DataFrame({"key": key, "val": val}).groupby("key")["val"].agg(["sum", "count", "mean", "min", "max"])
    assert np.array_equal(p.index.to_numpy(), gk[0][0])
    for name, (g, _) in zip(["sum", "count", "mean", "min", "max"], gr[0]):
        w = p[name].to_numpy()
        if name == "count":
            assert np.array_equal(g.astype(np.int64), w)
        else:
            assert np.allclose(g, w, rtol=1e-12, atol=0), name


def _c2_host(n, distinct, seed):
    rng = np.random.default_rng(seed)
    uniq = np.unique(rng.integers(-2 ** 63, 2 ** 63 - 1, distinct, dtype=np.int64))
    keys = uniq[rng.integers(0, uniq.size, n)]
    reqs = [((O.INT32, rng.integers(-1_000_000, 1_000_001, n).astype(np.int32), rng.random(n) >= 0.05), KINDS),
            ((O.FLOAT64, rng.standard_normal(n), rng.random(n) >= 0.05), KINDS),
            ((O.FLOAT32, rng.standard_normal(n, dtype=np.float32), rng.random(n) >= 0.05), KINDS)]
    return keys, reqs, uniq.size


@pytest.mark.timeout(900)
def test_c2_headline_workload_full_size(gpu):
    """The 100M-row workload at full size: partition count, work-item
    splitting and spill sizing at 1M groups, against the oracle."""
    t0 = time.time()
    keys, reqs, distinct = _c2_host(100_000_000, 1_000_000, 42)
    progress(f"c2 data generated in {time.time() - t0:.1f} s")
    gk, _ = _compare_to_oracle([(O.INT64, keys, None)], reqs, [None, None, 1e-5])
    assert len(gk[0][0]) == distinct
    progress(f"c2 full size done in {time.time() - t0:.1f} s")


def _device_abs_sums(keys_table, vals):
    """Per group (in the order of its own output) of a plain call: SUM of |x|."""
    av = ca.Column.from_torch(vals.data.view(torch.float64).abs(), ca.TypeId.FLOAT64, vals.mask, vals.null_count)
    k, r = ca.GroupBy(keys_table).aggregate([ca.GroupByRequest(av, [agg.sum()])])
    return k.columns[0].to_numpy()[0], r[0].columns[0].to_numpy()[0]


@pytest.mark.timeout(900)
def test_c4_batch_of_4096_queries(gpu):
    """4096 queries drawn like the batched workload, one batch (autotuned
    execution per query, one graph): each query equals its own plain call
    (keys and counts exact, float sums within tolerance), its group count is
    its number of distinct keys, and eight queries match the oracle."""
    queries, rows, card = bench.make_c4(4096, 42)
    progress(f"c4: {len(queries)} queries, {int(rows.sum()):,} rows")
    batch = ca.GroupByBatch(queries)
    st = batch.stats()
    progress(f"c4 batch stats {st}")
    batch.run()
    total_groups = 0
    sample = set(range(0, 4096, 512))
    for q, (kt, reqs) in enumerate(queries):
        bk, br = batch.result(q)
        bkeys = bk.columns[0].to_numpy()[0]
        bres = [c.to_numpy() for c in br[0].columns]
        pk, pr = ca.GroupBy(kt).aggregate(reqs)
        pkeys = pk.columns[0].to_numpy()[0]
        pres = [c.to_numpy() for c in pr[0].columns]
        distinct = int(torch.unique(kt.columns[0].data.view(torch.int32)).numel())
        assert bkeys.size == pkeys.size == distinct, f"query {q}: {bkeys.size} / {pkeys.size} groups, {distinct} keys"
        total_groups += bkeys.size
        bo, po = np.argsort(bkeys), np.argsort(pkeys)
        assert np.array_equal(bkeys[bo], pkeys[po]), f"query {q}: keys"
        ak, asum = _device_abs_sums(kt, reqs[0].values)
        s = asum[np.argsort(ak)]
        tol = [s, None, s / np.maximum(bres[1][0][bo], 1)]  # sum, count, mean
        for a, name in enumerate(["sum", "count", "mean"]):
            assert_same((bres[a][0][bo], None), (pres[a][0][po], None), None, abs_sum=tol[a],
                        what=f"query {q} {name} batch vs plain")
        if q in sample:
            hk = kt.columns[0].to_numpy()[0]
            hv = reqs[0].values.to_numpy()[0]
            ok, orr = oracle_groupby([(O.INT32, hk, None)], [((O.FLOAT64, hv, None), [O.SUM, O.COUNT_VALID, O.MEAN])])
            ok, orr = sorted_result(ok, orr)
            assert np.array_equal(ok[0][0], bkeys[bo]), f"query {q}: keys vs oracle"
            for a in range(3):
                assert_same((bres[a][0][bo], None), orr[0][a], None, abs_sum=tol[a], what=f"query {q} agg {a} vs oracle")
        if q % 512 == 511:
            progress(f"c4: {q + 1} queries checked")
    progress(f"c4: {total_groups:,} groups in all")


def _served_tol(r, a, abs_sum, results):
    """The abs-sum scale of request r's agg a (sum, count, min, max, mean):
    float64 SUM by the group's sum of |x|, MEAN by its mean of |x|; counts,
    MIN / MAX and the int32 column exact."""
    if r != 0 or a in (1, 2, 3):
        return None
    if a == 0:
        return abs_sum
    count, cvalid = results[0][1]
    return abs_sum / np.maximum(count, 1)


@pytest.mark.timeout(900)
def test_c5_served_mix_large_tables(gpu):
    """A continuously batched server over tables drawn like the serving
    workload, forced to hold two large ones (3e7 rows / 1e6 groups and 1e7
    rows / 2e5 groups): every query equals the same query run alone, and the
    largest table's result matches the oracle."""
    tables, rows, card = bench.make_c5_tables(10, 7, max_rows=2e6)
    big = []
    for n, c, seed in ((30_000_000, 1_000_000, 1), (10_000_000, 200_000, 2)):
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        pool = torch.randint(-2 ** 63, 2 ** 63 - 1, (c,), generator=g, device="cuda", dtype=torch.int64)
        keys = pool[torch.randint(0, c, (n,), generator=g, device="cuda")]
        f64 = torch.randn(n, generator=g, device="cuda", dtype=torch.float64)
        i32 = torch.randint(-1_000_000, 1_000_001, (n,), generator=g, device="cuda", dtype=torch.int32)
        cols = [ca.Column.from_torch(keys, ca.TypeId.INT64)]
        for v, tid in ((f64, ca.TypeId.FLOAT64), (i32, ca.TypeId.INT32)):
            valid = torch.rand(n, generator=g, device="cuda") >= 0.02
            cols.append(ca.Column.from_torch(v, tid, bench.pack_bits(valid), int((~valid).sum().item())))
        big.append(cols)
    tables = big + tables
    aggs = [agg.sum(), agg.count(), agg.min(), agg.max(), agg.mean()]
    srv = ca.Server(max_inflight=2, max_batch_queries=64)
    tids = [srv.add_table(t) for t in tables]
    rng = np.random.default_rng(8)
    picks = [0, 1] + [int(x) for x in rng.integers(0, len(tids), 46)]
    tickets = []
    for i, t in enumerate(picks):
        tickets.append(srv.submit(tids[t], [0], [(1, aggs), (2, aggs)]))
        time.sleep(float(rng.exponential(0.002)))
    got = {}
    for tk in tickets:
        srv.wait(tk, 300_000)
        k, r = srv.take(tk)
        got[tk] = ([c.to_numpy() for c in k.columns], [[c.to_numpy() for c in t.columns] for t in r])
    st = srv.stats()
    progress(f"c5 server stats {st}")
    assert st["completed"] == len(tickets) and st["failed"] == 0
    alone = {}
    for tk, t in zip(tickets, picks):
        if t not in alone:
            cols = tables[t]
            k, r = ca.GroupBy(ca.Table([cols[0]])).aggregate(
                [ca.GroupByRequest(cols[1], aggs), ca.GroupByRequest(cols[2], aggs)])
            ak, asum = _device_abs_sums(ca.Table([cols[0]]), cols[1])
            alone[t] = ([c.to_numpy() for c in k.columns], [[c.to_numpy() for c in tt.columns] for tt in r],
                        asum[np.argsort(ak)])
        wk, wr, s = alone[t]
        gk, gr = sorted_result(*got[tk])
        wk2, wr2 = sorted_result(wk, wr)
        assert np.array_equal(gk[0][0], wk2[0][0]), f"ticket {tk} (table {t}): keys"
        for r in range(2):
            for a in range(5):
                assert_same(gr[r][a], wr2[r][a], None, abs_sum=_served_tol(r, a, s, wr2),
                            what=f"ticket {tk} table {t} request {r} agg {a} mix vs alone")
    # the largest table against the oracle
    cols = tables[0]
    hk = cols[0].to_numpy()[0]
    (h64, v64), (h32, v32) = cols[1].to_numpy(), cols[2].to_numpy()
    progress("c5: oracle on the 3e7-row table")
    ok, orr = oracle_groupby([(O.INT64, hk, None)], [((O.FLOAT64, h64, v64), KINDS), ((O.INT32, h32, v32), KINDS)])
    ok, orr = sorted_result(ok, orr)
    wk, wr, s = alone[0]
    wk2, wr2 = sorted_result(wk, wr)
    assert np.array_equal(ok[0][0], wk2[0][0])
    # the served kinds are in the order sum, count, min, max, mean; the oracle's in KINDS order
    order = [KINDS.index(k) for k in (O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN)]
    for r in range(2):
        for a in range(5):
            assert_same(wr2[r][a], orr[r][order[a]], None, abs_sum=_served_tol(r, a, s, wr2),
                        what=f"large table request {r} agg {a} vs oracle")
    srv.close()
