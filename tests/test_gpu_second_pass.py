"""GPU parity of the second-pass aggregations (VARIANCE, STD, M2, ARGMIN, ARGMAX)
through the C ABI: against the oracle on seeded inputs (every value type,
nulls, ties, NaN and signed zeros, composite and null keys, several ddof),
size-independent properties at 20M rows, and their routing (batches refuse
them; a batching server runs them alone on the plain path).

Tolerances: ARGMIN / ARGMAX are row indices and must match exactly.
VARIANCE is compared as |gpu - oracle| <= rtol * (sum of x^2) / (n - ddof)
per group (M2 against rtol * sum of x^2), rtol 1e-12 (1e-9 for float32 sources, whose mean comes from a
float32 SUM): the deviation sum has no cancellation, so its error is relative
to the second moment; STD takes the square root of that bound."""
from __future__ import annotations

import numpy as np
import pytest

import cudf_amd as ca
from cudf_amd import aggregation as agg
from oracle import oracle as O
from tests.helpers import abs_sums, assert_same, gpu_groupby, oracle_groupby, sorted_result

pytestmark = pytest.mark.gpu

SECOND = [(O.VARIANCE, 1), (O.STD, 1), O.M2, O.ARGMIN, O.ARGMAX]


def _kind(a):
    return a[0] if isinstance(a, tuple) else a


def _moments(keys, reqs, include_null_keys):
    """Per request: (count of valid values, sum of x^2) per group, sorted by key."""
    out = []
    for (t, v, valid), _ in reqs:
        x = np.asarray(v, dtype=np.float64)
        x = np.where(np.isfinite(x), x, 0.0)
        kcols = [O.HostColumn(kt, kv, kvalid) for kt, kv, kvalid in keys]
        ks, res = O.groupby(kcols, [(O.HostColumn(O.FLOAT64, x * x, valid), [O.SUM, O.COUNT_VALID])],
                            include_null_keys)
        order = O.sort_order(ks)
        out.append((res[0][1][0][order].astype(np.float64), res[0][0][0][order]))
    return out


def _compare(keys, reqs, include_null_keys=False):
    gk, gr = gpu_groupby(keys, reqs, include_null_keys)
    ok, orr = oracle_groupby(keys, reqs, include_null_keys)
    gk, gr = sorted_result(gk, gr)
    ok, orr = sorted_result(ok, orr)
    assert len(gk[0][0]) == len(ok[0][0]), "group count differs"
    for (gv, gvalid), (ov, ovalid) in zip(gk, ok):
        gm = np.ones(len(gv), bool) if gvalid is None else np.asarray(gvalid, bool)
        om = np.ones(len(ov), bool) if ovalid is None else np.asarray(ovalid, bool)
        assert np.array_equal(gm, om) and np.array_equal(gv[om], ov[om]), "keys differ"
    moments = _moments(keys, reqs, include_null_keys)
    sums = abs_sums(keys, reqs, include_null_keys)
    for r, (gblk, oblk) in enumerate(zip(gr, orr)):
        cnt, sumsq = moments[r]
        sk, s = sums[r]
        f32 = reqs[r][0][0] == O.FLOAT32
        for a, ((g, gvalid), (o, ovalid)) in zip(reqs[r][1], zip(gblk, oblk)):
            what = f"request {r} agg {a}"
            if _kind(a) not in (O.VARIANCE, O.STD, O.M2, O.ARGMIN, O.ARGMAX):
                # the first-pass aggregations, to the bar of the other parity tests
                assert_same((g, gvalid), (o, ovalid), None, abs_sum=s[O.sort_order(sk)],
                            rtol=1e-5 if f32 else None, what=what)
                continue
            gm = np.ones(len(g), bool) if gvalid is None else np.asarray(gvalid, bool)
            om = np.ones(len(o), bool) if ovalid is None else np.asarray(ovalid, bool)
            assert np.array_equal(gm, om), f"{what}: validity differs at {np.flatnonzero(gm != om)[:5]}"
            g, o = np.asarray(g)[om], np.asarray(o)[om]
            k = _kind(a)
            if k in (O.VARIANCE, O.STD, O.M2):
                ddof = a[1] if isinstance(a, tuple) else 1
                rtol = 1e-9 if f32 else 1e-12
                scale = sumsq[om] if k == O.M2 else sumsq[om] / np.maximum(cnt[om] - ddof, 1)
                bound = rtol * scale if k == O.VARIANCE else np.sqrt(rtol * scale)
                err = np.abs(g - o)
                bad = ~((err <= bound) | (err <= 1e-12 * np.abs(o)))
                assert not bad.any(), f"{what}: {int(bad.sum())} groups outside tolerance, worst {err[bad].max()}"
            else:
                assert np.array_equal(g, o), f"{what}: indices differ at {np.flatnonzero(g != o)[:5]}"


def _values(rng, vt, n, ties=False):
    if vt in (O.FLOAT32, O.FLOAT64):
        if ties:
            pool = np.array([-np.inf, -1.5, -0.0, 0.0, 2.5, np.inf, np.nan], dtype=O.NP[vt])
            return pool[rng.integers(0, len(pool), n)]
        return (rng.standard_normal(n) * 50 + 3).astype(O.NP[vt])
    if vt == O.BOOL8:
        return rng.integers(0, 2, n).astype(np.uint8)
    info = np.iinfo(O.NP[vt])
    if ties:
        pool = np.array([info.min, 0, 1, info.max], dtype=O.NP[vt])
        return pool[rng.integers(0, len(pool), n)]
    return rng.integers(max(info.min, -10 ** 6), min(info.max, 10 ** 6), n, endpoint=True).astype(O.NP[vt])


@pytest.mark.parametrize("vt", [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT8, O.UINT32, O.UINT64, O.FLOAT32,
                                O.FLOAT64, O.BOOL8])
def test_value_types(gpu, vt):
    rng = np.random.default_rng(500 + vt)
    n = 60_000
    keys = rng.integers(-2 ** 31, 2 ** 31 - 1, 4000, dtype=np.int64).astype(np.int32)[rng.integers(0, 4000, n)]
    reqs = [((vt, _values(rng, vt, n), rng.random(n) > 0.1), SECOND + [O.SUM, O.MEAN])]
    _compare([(O.INT32, keys, None)], reqs)


@pytest.mark.parametrize("vt", [O.INT32, O.UINT64, O.FLOAT32, O.FLOAT64])
def test_ties_nan_and_signed_zero(gpu, vt):
    """Extreme values repeated within groups: the first row holding one is
    the result; NaN only where a group's valid values are all NaN."""
    rng = np.random.default_rng(520 + vt)
    n = 80_000
    keys = rng.integers(0, 900, n).astype(np.int64)
    vals = _values(rng, vt, n, ties=True)
    if vt in (O.FLOAT32, O.FLOAT64):
        vals[keys == 7] = np.nan
    _compare([(O.INT64, keys, None)], [((vt, vals, rng.random(n) > 0.05), [O.ARGMIN, O.ARGMAX, O.MIN, O.MAX])])


@pytest.mark.parametrize("ddof", [0, 1, 2, 5])
def test_ddof_and_small_groups(gpu, ddof):
    rng = np.random.default_rng(540 + ddof)
    n = 30_000
    # group sizes from 1 to ~10: many groups at or under ddof
    keys = rng.integers(0, n // 4, n).astype(np.int32)
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.3), [(O.VARIANCE, ddof), (O.STD, ddof)]),
            ((O.INT64, _values(rng, O.INT64, n), None), [(O.VARIANCE, ddof), (O.STD, ddof), (O.VARIANCE, 1)])]
    _compare([(O.INT32, keys, None)], reqs)


def test_composite_and_null_keys(gpu):
    rng = np.random.default_rng(560)
    n = 100_000
    a = rng.integers(0, 200, n).astype(np.int32)
    b = rng.integers(0, 40, n).astype(np.int64) * 1_000_003
    keys = [(O.INT32, a, rng.random(n) > 0.05), (O.INT64, b, None)]
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.1), SECOND),
            ((O.INT16, _values(rng, O.INT16, n), None), [O.ARGMAX, (O.STD, 0)])]
    _compare(keys, reqs)
    _compare(keys, reqs, include_null_keys=True)


PATHS = ["auto", "partitioned", "direct", "global"]


@pytest.fixture(params=PATHS)
def path(request, monkeypatch):
    """Each execution runs the second pass its own way: the partitioned items
    in LDS after their first pass (split or spilled items on the global
    table), the direct workgroups on LDS copies of the completed table, the
    global execution on the global table."""
    if request.param != "auto":
        monkeypatch.setenv("MGB_GROUPBY_PATH", request.param)
    else:
        monkeypatch.delenv("MGB_GROUPBY_PATH", raising=False)
    return request.param


def test_every_execution(gpu, path):
    rng = np.random.default_rng(580)
    n = 200_000
    keys = rng.integers(0, 50_000, n).astype(np.int64)
    reqs = [((O.FLOAT32, _values(rng, O.FLOAT32, n), rng.random(n) > 0.05), SECOND + [O.COUNT_VALID]),
            ((O.INT32, _values(rng, O.INT32, n, ties=True), rng.random(n) > 0.05), [O.ARGMAX, (O.VARIANCE, 0)])]
    _compare([(O.INT64, keys, None)], reqs)


@pytest.mark.parametrize("rows,card", [(3_000_000, 40), (2_000_000, 1000), (400_000, 120_000)])
def test_few_and_many_groups(gpu, path, rows, card):
    # few groups over many rows: partitions split over several items (their
    # second pass on the global table) or the direct execution; many groups:
    # more than one LDS table holds
    rng = np.random.default_rng(590 + card)
    keys = rng.integers(-10 ** 9, 10 ** 9, card)[rng.integers(0, card, rows)].astype(np.int32)
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, rows), rng.random(rows) > 0.02), SECOND),
            ((O.INT16, _values(rng, O.INT16, rows, ties=True), None), [O.ARGMIN, O.ARGMAX])]
    _compare([(O.INT32, keys, None)], reqs)


def test_row_keys_in_every_execution(gpu, path):
    # float and composite keys (the row-reference table headers)
    rng = np.random.default_rng(595)
    n = 150_000
    a = rng.integers(0, 300, n).astype(np.float64) / 4
    b = rng.integers(0, 30, n).astype(np.int16)
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.1), SECOND)]
    _compare([(O.FLOAT64, a, None), (O.INT16, b, None)], reqs)
    _compare([(O.FLOAT64, a, None)], reqs)


def test_empty_and_all_null(gpu):
    gk, gr = gpu_groupby([(O.INT32, np.zeros(0, np.int32), None)], [((O.FLOAT64, np.zeros(0), None), SECOND)])
    assert len(gk[0][0]) == 0 and all(len(v) == 0 for v, _ in gr[0])
    keys = np.array([5, 5, 6], np.int32)
    vals = np.array([1.0, 2.0, 3.0])
    _compare([(O.INT32, keys, None)], [((O.FLOAT64, vals, np.array([False, False, False])), SECOND)])
    _compare([(O.INT32, keys, np.array([False, False, False]))], [((O.FLOAT64, vals, None), SECOND)])


def test_properties_at_20m_rows(gpu):
    """At 20M rows over 1M groups: vals[ARGMAX] is MAX and vals[ARGMIN] is MIN
    for every group, the rows' keys are the group's key, no earlier row of
    the group holds the extreme (checked on a sample of groups), and
    VARIANCE agrees with SUM_OF_SQUARES - SUM^2 / n."""
    rng = np.random.default_rng(600)
    n, card = 20_000_000, 1_000_000
    keys = rng.integers(0, card, n, dtype=np.int64) * 7919
    vals = rng.integers(-1000, 1000, n).astype(np.int32)  # many ties per group
    valid = rng.random(n) > 0.05
    gk, gr = gpu_groupby([(O.INT64, keys, None)],
                         [((O.INT32, vals, valid), [O.ARGMIN, O.ARGMAX, O.MIN, O.MAX, (O.VARIANCE, 1),
                                                    O.SUM, O.SUM_OF_SQUARES, O.COUNT_VALID])])
    k = gk[0][0]

    def mask(v, m):
        return np.ones(len(v), bool) if m is None else np.asarray(m, bool)

    (amin, amin_ok), (amax, amax_ok), (mn, _), (mx, _), (var, var_ok), (s, _), (s2, _), (c, _) = gr[0]
    assert len(k) == len(np.unique(keys))
    ok = mask(amin, amin_ok)
    assert np.array_equal(ok, mask(amax, amax_ok)) and np.array_equal(ok, c > 0)
    assert np.array_equal(keys[amin[ok]], k[ok]) and np.array_equal(keys[amax[ok]], k[ok])
    assert valid[amin[ok]].all() and valid[amax[ok]].all()
    assert np.array_equal(vals[amin[ok]], mn[ok]) and np.array_equal(vals[amax[ok]], mx[ok])
    # first holder of the extreme: no valid row of the group before it holds it
    pick = rng.choice(np.flatnonzero(ok), 50, replace=False)
    for g in pick:
        rows = np.flatnonzero((keys == k[g]) & valid)
        assert rows[vals[rows] == mx[g]].min() == amax[g] and rows[vals[rows] == mn[g]].min() == amin[g]
    vv = mask(var, var_ok)
    assert np.array_equal(vv, c > 1)
    cf = c[vv].astype(np.float64)
    ref = (s2[vv].astype(np.float64) - s[vv].astype(np.float64) ** 2 / cf) / (cf - 1)
    assert np.allclose(var[vv], ref, rtol=1e-9, atol=1e-9)


def test_batch_refuses_second_pass(gpu):
    kc = ca.Column.from_numpy(np.arange(100, dtype=np.int32) % 7)
    vc = ca.Column.from_numpy(np.arange(100, dtype=np.float64))
    with pytest.raises(ca._lib.LogicError, match="second pass"):
        ca.GroupByBatch([(ca.Table([kc]), [ca.GroupByRequest(vc, [agg.variance()])])])


def test_server_runs_second_pass_alone(gpu):
    rng = np.random.default_rng(620)
    n = 50_000
    keys = rng.integers(0, 3000, n).astype(np.int64)
    v = rng.standard_normal(n)
    valid = rng.random(n) > 0.05
    cols = [ca.Column.from_numpy(keys, None, ca.TypeId.INT64), ca.Column.from_numpy(v, valid, ca.TypeId.FLOAT64)]
    aggs = [agg.sum(), agg.variance(), agg.std(ddof=0), agg.argmin(), agg.argmax()]
    srv = ca.Server(max_inflight=2, max_batch_queries=8)
    tid = srv.add_table(cols)
    tickets = [srv.submit(tid, [0], [(1, aggs)]) for _ in range(6)]
    plain = [srv.submit(tid, [0], [(1, [agg.sum(), agg.mean()])]) for _ in range(6)]
    want_k, want_r = oracle_groupby([(O.INT64, keys, None)],
                                    [((O.FLOAT64, v, valid), [O.SUM, (O.VARIANCE, 1), (O.STD, 0), O.ARGMIN,
                                                             O.ARGMAX])])
    want_k, want_r = sorted_result(want_k, want_r)
    for tk in tickets:
        srv.wait(tk, 60_000)
        kt, rt = srv.take(tk)
        gk, gr = sorted_result([c.to_numpy() for c in kt.columns], [[c.to_numpy() for c in t.columns] for t in rt])
        assert np.array_equal(gk[0][0], want_k[0][0])
        for (g, gv), (o, ov) in zip(gr[0], want_r[0]):
            gm = np.ones(len(g), bool) if gv is None else np.asarray(gv, bool)
            om = np.ones(len(o), bool) if ov is None else np.asarray(ov, bool)
            assert np.array_equal(gm, om)
            if g.dtype.kind == "f":
                assert np.allclose(g[om], o[om], rtol=1e-10, atol=1e-12)
            else:
                assert np.array_equal(g[om], o[om])
    for tk in plain:
        srv.wait(tk, 60_000)
        srv.release(tk)
    st = srv.stats()
    assert st["completed"] == 12 and st["failed"] == 0, st
    srv.close()
