"""GPU tests of serving (mgb_server_* through the C ABI): every query's
result is the same when it runs alone (the plain path, one call) and when it
runs in the mix of a continuously batched server, within the stated
tolerance (integers exact; float64 rel 1e-12, float32 rel 1e-5, or the same
scaled by the group's sum of |x|); HBM-budget admission; the per-query
baseline mode; errors."""
from __future__ import annotations

import threading
import time

import numpy as np
import pytest

import cudf_amd as ca
from cudf_amd import aggregation as agg
from oracle import oracle as O
from tests.helpers import abs_sums, assert_same, oracle_groupby, sorted_result

pytestmark = pytest.mark.gpu

AGGS = [agg.sum(), agg.count(), agg.min(), agg.max(), agg.mean()]
KINDS = [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN]


def _tables(rng, n_tables, max_rows=200_000):
    """Host copies and device columns of tables: int64 key, f64 and int32 values, 2% nulls."""
    host, dev = [], []
    for _ in range(n_tables):
        rows = int(10 ** rng.uniform(3, np.log10(max_rows)))
        card = int(min(rows, 10 ** rng.uniform(1, 5)))
        pool = rng.integers(-2 ** 62, 2 ** 62, card, dtype=np.int64)
        keys = pool[rng.integers(0, card, rows)]
        f64 = rng.standard_normal(rows)
        i32 = rng.integers(-1_000_000, 1_000_000, rows, endpoint=True).astype(np.int32)
        v64, v32 = rng.random(rows) >= 0.02, rng.random(rows) >= 0.02
        host.append((keys, (f64, v64), (i32, v32)))
        dev.append([ca.Column.from_numpy(keys, None, ca.TypeId.INT64),
                    ca.Column.from_numpy(f64, v64, ca.TypeId.FLOAT64),
                    ca.Column.from_numpy(i32, v32, ca.TypeId.INT32)])
    return host, dev


def _numpy(keys_t, results):
    return [c.to_numpy() for c in keys_t.columns], [[c.to_numpy() for c in t.columns] for t in results]


def _same(got, want, host_table, what):
    """got / want: (keys, results) as numpy; compared after sorting by key."""
    gk, gr = sorted_result(*got)
    wk, wr = sorted_result(*want)
    assert len(gk[0][0]) == len(wk[0][0]), f"{what}: group count"
    assert_same(gk[0], wk[0], None, what=f"{what} keys")
    keys, (f64, v64), (i32, v32) = host_table
    kspec = [(O.INT64, keys, None)]
    sums = abs_sums(kspec, [((O.FLOAT64, f64, v64), KINDS), ((O.INT32, i32, v32), KINDS)])
    for r, (gblk, wblk) in enumerate(zip(gr, wr)):
        sk, s = sums[r]
        s_sorted = s[O.sort_order(sk)]
        for a, (g, w) in enumerate(zip(gblk, wblk)):
            assert_same(g, w, None, abs_sum=s_sorted, what=f"{what} request {r} agg {KINDS[a]}")


def _alone(dev_table):
    gb = ca.GroupBy(ca.Table([dev_table[0]]))
    k, r = gb.aggregate([ca.GroupByRequest(dev_table[1], AGGS), ca.GroupByRequest(dev_table[2], AGGS)])
    return _numpy(k, r)


def _submit(srv, tid):
    return srv.submit(tid, [0], [(1, AGGS), (2, AGGS)])


def _drain(srv, tickets, timeout_s=60.0):
    """Takes every ticket's result in completion order (held results count
    against the server's budget, so a client that waits in submission order
    can hold memory a waiting query needs)."""
    out, left, t0 = {}, list(tickets), time.time()
    while left:
        assert time.time() - t0 < timeout_s, f"{len(left)} tickets unfinished"
        for tk in list(left):
            try:
                srv.wait(tk, 1.0)
            except ca._lib.LogicError as e:
                if "timed out" in str(e):
                    continue
                raise
            out[tk] = _numpy(*srv.take(tk))
            left.remove(tk)
    return out


def test_result_same_alone_and_in_the_mix(gpu):
    rng = np.random.default_rng(21)
    host, dev = _tables(rng, 8)
    srv = ca.Server(max_inflight=2, max_batch_queries=16)
    tids = [srv.add_table(t) for t in dev]
    picks = rng.integers(0, len(tids), 48)
    # arrivals spread over time so that batches of different composition form
    # while others run
    tickets = []
    for i, t in enumerate(picks):
        tickets.append(_submit(srv, tids[t]))
        if i % 6 == 5:
            time.sleep(0.002)
    mixed = []
    for tk in tickets:
        srv.wait(tk, 60_000)
        mixed.append(_numpy(*srv.take(tk)))
    st = srv.stats()
    print("server stats", st, flush=True)
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["batches"] < len(tickets), "queries were not batched"
    alone = {}
    for i, t in enumerate(picks):
        if t not in alone:
            alone[t] = _alone(dev[t])
        _same(mixed[i], alone[t], host[t], f"query {i} (table {t})")
    # and the alone result against the oracle, for two tables
    for t in list(alone)[:2]:
        keys, (f64, v64), (i32, v32) = host[t]
        want = oracle_groupby([(O.INT64, keys, None)], [((O.FLOAT64, f64, v64), KINDS), ((O.INT32, i32, v32), KINDS)])
        _same(alone[t], want, host[t], f"table {t} alone vs oracle")
    srv.close()


def test_concurrent_submitters(gpu):
    rng = np.random.default_rng(22)
    host, dev = _tables(rng, 4, max_rows=50_000)
    srv = ca.Server(max_inflight=3, max_batch_queries=8)
    tids = [srv.add_table(t) for t in dev]
    tickets, lock = [], threading.Lock()

    def client(seed):
        r = np.random.default_rng(seed)
        for _ in range(10):
            t = int(r.integers(0, len(tids)))
            tk = _submit(srv, tids[t])
            with lock:
                tickets.append((tk, t))

    threads = [threading.Thread(target=client, args=(s,)) for s in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    alone = {t: _alone(dev[t]) for t in range(len(dev))}
    for tk, t in tickets:
        srv.wait(tk, 60_000)
        _same(_numpy(*srv.take(tk)), alone[t], host[t], f"ticket {tk}")
    srv.close()


def test_hbm_budget_admission(gpu):
    rng = np.random.default_rng(23)
    host, dev = _tables(rng, 4, max_rows=100_000)
    resident = sum(sum(c.data.numel() * c.data.element_size() + (c.mask.numel() * 4 if c.mask is not None else 0)
                       for c in t) for t in dev)
    # device memory of the largest single-query batch: the budget leaves room
    # for the tables and about one query at a time
    one = max(ca.GroupByBatch([(ca.Table([t[0]]), [ca.GroupByRequest(t[1], AGGS), ca.GroupByRequest(t[2], AGGS)])])
              .stats()["device_bytes"] for t in dev)
    budget = resident + int(one * 1.5)
    print("resident", resident, "single-query batch bytes", one, "budget", budget, flush=True)
    srv = ca.Server(hbm_budget_bytes=budget, max_inflight=2, max_batch_queries=32)
    tids = [srv.add_table(t) for t in dev]
    picks = rng.integers(0, len(tids), 24)
    tickets = [_submit(srv, tids[t]) for t in picks]
    got = _drain(srv, tickets)
    for tk, t in zip(tickets, picks):
        _same(got[tk], _alone(dev[t]), host[t], f"ticket {tk}")
    st = srv.stats()
    print("budget stats", st, flush=True)
    assert st["completed"] == len(tickets) and st["failed"] == 0
    assert st["deferred"] > 0, st
    assert st["peak_held_bytes"] + st["resident_bytes"] <= budget, st
    srv.close()


def test_reserved_memory(gpu):
    rng = np.random.default_rng(26)
    host, dev = _tables(rng, 3, max_rows=100_000)
    with pytest.raises(ca._lib.AllocError):
        ca.Server(hbm_budget_bytes=64 << 20, reserve_bytes=128 << 20)
    srv = ca.Server(reserve_bytes=256 << 20, max_inflight=2, max_batch_queries=4)
    assert srv.stats()["device_bytes"] == 256 << 20
    tids = [srv.add_table(t) for t in dev]
    picks = rng.integers(0, len(tids), 16)
    tickets = [_submit(srv, tids[t]) for t in picks]
    got = _drain(srv, tickets)
    for tk, t in zip(tickets, picks):
        _same(got[tk], _alone(dev[t]), host[t], f"ticket {tk}")
    st = srv.stats()
    print("reserve stats", st, flush=True)
    # every batch was carved from the reserved chunk: nothing more was taken
    assert st["device_bytes"] == 256 << 20 and st["held_bytes"] == 0, st
    srv.close()


def test_baseline_mode_matches(gpu):
    rng = np.random.default_rng(24)
    host, dev = _tables(rng, 3, max_rows=50_000)
    srv = ca.Server(batching=False)
    tids = [srv.add_table(t) for t in dev]
    tickets = [(_submit(srv, tids[t]), t) for t in (0, 1, 2, 1)]
    for tk, t in tickets:
        srv.wait(tk, 60_000)
        _same(_numpy(*srv.take(tk)), _alone(dev[t]), host[t], f"ticket {tk}")
    assert srv.stats()["batches"] == 4
    srv.close()


def test_errors(gpu):
    rng = np.random.default_rng(25)
    _, dev = _tables(rng, 1, max_rows=2000)
    srv = ca.Server(hbm_budget_bytes=4 << 20)
    with pytest.raises(ca._lib.AllocError):
        big = [ca.Column.from_numpy(np.zeros(1_000_000, np.int64), None, ca.TypeId.INT64)]
        srv.add_table(big)
    tid = srv.add_table(dev[0])
    with pytest.raises(ca._lib.LogicError):
        srv.submit(tid + 5, [0], [(1, AGGS)])
    with pytest.raises(ca._lib.LogicError):
        srv.submit(tid, [7], [(1, AGGS)])
    tk = _submit(srv, tid)
    srv.wait(tk, 60_000)
    srv.release(tk)
    with pytest.raises(ca._lib.LogicError):
        srv.release(tk)
    srv.close()
