"""GPU parity of the batched execution (mgb_batch_* through the C ABI): every
query of a batch is compared with the oracle run on that query alone, for
each execution the planner can pick (forced and automatic), with and without
graph capture, across repeated executions, and through the plain-path re-run
that follows an underestimated group count."""
from __future__ import annotations

import numpy as np
import pytest

import cudf_amd as ca
from oracle import oracle as O
from tests.helpers import abs_sums, assert_same, oracle_groupby, sorted_result, to_agg

pytestmark = pytest.mark.gpu

KINDS = [O.SUM, O.COUNT_VALID, O.MEAN, O.MIN, O.MAX]


def _query(rng, rows, card, key_type=O.INT32, null_frac=0.0, val_types=(O.FLOAT64,), kinds=KINDS):
    info = np.iinfo(O.NP[key_type])
    pool = rng.choice(np.arange(max(info.min, -2 ** 40), max(info.min, -2 ** 40) + 4 * card + 16), card,
                      replace=False).astype(O.NP[key_type]) if key_type != O.INT32 else \
        rng.integers(info.min, info.max, card, dtype=np.int64).astype(np.int32)
    keys = pool[rng.integers(0, card, rows)]
    reqs = []
    for vt in val_types:
        if vt in (O.FLOAT32, O.FLOAT64):
            v = rng.standard_normal(rows).astype(O.NP[vt])
        else:
            v = rng.integers(-1_000_000, 1_000_000, rows, endpoint=True).astype(O.NP[vt])
        valid = (rng.random(rows) >= null_frac) if null_frac > 0 else None
        reqs.append(((vt, v, valid), list(kinds)))
    return [(key_type, keys, None)], reqs


def _device_query(keys, reqs):
    kcols = [ca.Column.from_numpy(np.asarray(v), valid, t) for t, v, valid in keys]
    dreqs = [ca.GroupByRequest(ca.Column.from_numpy(np.asarray(v), valid, t), [to_agg(k) for k in kinds])
             for (t, v, valid), kinds in reqs]
    return ca.Table(kcols), dreqs


def _check(batch, q, keys, reqs):
    gk_t, gr_t = batch.result(q)
    gk = [c.to_numpy() for c in gk_t.columns]
    gr = [[c.to_numpy() for c in t.columns] for t in gr_t]
    ok, orr = oracle_groupby(keys, reqs)
    gk, gr = sorted_result(gk, gr)
    ok, orr = sorted_result(ok, orr)
    assert len(gk[0][0]) == len(ok[0][0]), f"query {q}: {len(gk[0][0])} groups vs {len(ok[0][0])}"
    for (gv, gvalid), (ov, ovalid) in zip(gk, ok):
        assert_same((gv, gvalid), (ov, ovalid), None, what=f"query {q} keys")
    sums = abs_sums(keys, reqs)
    for r, (gblk, oblk) in enumerate(zip(gr, orr)):
        sk, s = sums[r]
        s_sorted = s[O.sort_order(sk)]
        rtol = 1e-5 if reqs[r][0][0] == O.FLOAT32 else None
        for a, (g, o) in enumerate(zip(gblk, oblk)):
            assert_same(g, o, None, abs_sum=s_sorted, rtol=rtol, what=f"query {q} request {r} agg {reqs[r][1][a]}")


# (rows, groups) per query: tiny, small-and-dense, wide, large-and-few, large-and-many
SHAPES = [(1, 1), (37, 5), (1000, 1000), (5000, 40), (20000, 3000), (120000, 17), (150000, 30000), (300000, 90000)]


def _batch_inputs(seed, shapes=SHAPES, **kw):
    rng = np.random.default_rng(seed)
    host = [_query(rng, r, c, **kw) for r, c in shapes]
    dev = [_device_query(k, q) for k, q in host]
    return host, dev


@pytest.mark.parametrize("mode", [None, 0, 1, 2], ids=["auto", "lds_direct", "global", "partitioned"])
def test_batch_modes(gpu, mode):
    host, dev = _batch_inputs(11)
    batch = ca.GroupByBatch(dev, force_mode=mode)
    print("batch modes", [batch.mode(q) for q in range(len(dev))], batch.stats(), flush=True)
    batch.run()
    for q, (keys, reqs) in enumerate(host):
        _check(batch, q, keys, reqs)


@pytest.mark.parametrize("mode", [None, 2], ids=["auto", "partitioned"])
def test_batch_long_partitions_are_split(gpu, mode):
    # few groups over many rows: partitions keep 64 groups or more, so there
    # are few of them, each cut into several work items whose groups meet in
    # the query's table
    host, dev = _batch_inputs(18, shapes=[(4_000_000, 1000), (3_000_000, 800), (40000, 40)],
                              val_types=(O.FLOAT64, O.INT32), null_frac=0.02)
    batch = ca.GroupByBatch(dev, force_mode=mode)
    print("split modes", [batch.mode(q) for q in range(len(dev))], batch.stats(), flush=True)
    batch.run()
    for q, (keys, reqs) in enumerate(host):
        _check(batch, q, keys, reqs)


def test_batch_auto_picks_by_shape(gpu):
    host, dev = _batch_inputs(12)
    batch = ca.GroupByBatch(dev)
    modes = [batch.mode(q)[0] for q in range(len(dev))]
    # few groups per row -> LDS direct; a large query with many groups -> partitioned
    assert modes[5] == "lds_direct"
    assert modes[-1] == "partitioned"
    batch.run()
    for q, (keys, reqs) in enumerate(host):
        _check(batch, q, keys, reqs)


def test_batch_nulls_and_types(gpu):
    host, dev = _batch_inputs(13, shapes=[(5000, 50), (80000, 20000), (200, 199), (60000, 6)],
                              key_type=O.INT64, null_frac=0.05, val_types=(O.INT32, O.FLOAT32, O.FLOAT64))
    for mode in (None, 2):
        batch = ca.GroupByBatch(dev, force_mode=mode)
        batch.run()
        for q, (keys, reqs) in enumerate(host):
            _check(batch, q, keys, reqs)


def test_batch_without_graph_and_repeated(gpu):
    host, dev = _batch_inputs(14)
    for use_graph in (True, False):
        batch = ca.GroupByBatch(dev, use_graph=use_graph)
        for _ in range(3):  # every execution starts from cleared tables and counters
            batch.run()
            for q, (keys, reqs) in enumerate(host):
                _check(batch, q, keys, reqs)


def test_batch_rerun_after_underestimate(gpu):
    # estimates scaled far down: global tables overflow and those queries are
    # re-run on the plain path by finish()
    host, dev = _batch_inputs(15, shapes=[(40000, 20000), (30000, 12000), (5000, 10)])
    batch = ca.GroupByBatch(dev, force_mode=0, estimate_scale=0.01)
    batch.run()
    assert batch.stats()["reruns"] >= 1
    for q, (keys, reqs) in enumerate(host):
        _check(batch, q, keys, reqs)


def test_batch_with_given_estimates(gpu):
    # estimates from mgb_estimate_groups (the cache a server keeps for resident
    # tables) replace the device sample
    import ctypes
    from cudf_amd import _lib
    host, dev = _batch_inputs(19)
    lib = _lib.load()
    pool = ca.DevicePool(1 << 30)
    ests = []
    for kt, _ in dev:
        arr = (_lib.mgb_column * 1)(ca.groupby._to_mgb(kt.columns[0]))
        e = ctypes.c_int64()
        _lib.check(lib.mgb_estimate_groups(arr, 1, 0, ctypes.c_void_p(0), ctypes.byref(pool.allocator),
                                           ctypes.byref(e)))
        ests.append(e.value)
    for (keys, _), e in zip(host, ests):
        exact = len(np.unique(keys[0][1]))
        assert abs(e - exact) <= max(2, 0.1 * exact), (e, exact)
    batch = ca.GroupByBatch(dev, estimates=ests)
    batch.run()
    for q, (keys, reqs) in enumerate(host):
        _check(batch, q, keys, reqs)


def test_batch_schema_mismatch_is_logic_error(gpu):
    rng = np.random.default_rng(16)
    k1, r1 = _query(rng, 100, 10)
    k2, r2 = _query(rng, 100, 10, val_types=(O.INT32,))
    with pytest.raises(ca._lib.LogicError):
        ca.GroupByBatch([_device_query(k1, r1), _device_query(k2, r2)])


def test_batch_many_small_queries(gpu):
    rng = np.random.default_rng(17)
    shapes = [(int(r), int(min(r, c))) for r, c in
              zip(10 ** rng.uniform(1, 4, 300), 10 ** rng.uniform(0, 3, 300))]
    host = [_query(rng, r, c, kinds=[O.SUM, O.COUNT_VALID, O.MEAN]) for r, c in shapes]
    dev = [_device_query(k, q) for k, q in host]
    batch = ca.GroupByBatch(dev)
    batch.run()
    for q in rng.choice(len(host), 40, replace=False):
        _check(batch, int(q), *host[int(q)])
