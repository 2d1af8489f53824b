"""GPU parity of COLLECT_LIST and COLLECT_SET through the C ABI: every
literal fixture of the reference's collect tests, then seeded inputs against
the oracle: every fixed-width value type, both null policies, both null and
NaN equalities, -0.0 / +0.0 and NaN payloads, composite and null keys, each
requested execution, empty and all-null inputs, several list outputs in one
call beside hash aggregations, and 2M rows. Lists must match the oracle's
element for element (row order); sets must equal the oracle's set sorted
ascending with NaN after numbers and nulls last, the order this library
returns them in."""
from __future__ import annotations

import numpy as np
import pytest

import cudf_amd as ca
from oracle import oracle as O
from tests import golden_cases as G

pytestmark = pytest.mark.gpu

COLLECT_CASES = G.load_collect_cases()


def _agg(kind, nulls=1, neq=0, naneq=0):
    if kind == O.COLLECT_LIST:
        return ca.aggregation.collect_list(ca.NullPolicy(nulls))
    return ca.aggregation.collect_set(ca.NullPolicy(nulls), ca.NullEquality(neq), ca.NanEquality(naneq))


def gpu_collect(keys, vals, aggs, include_null_keys=False, extra=()):
    """keys: [(type_id, values, valid)], vals: (type_id, values, valid), aggs:
    [(kind, nulls, nulls_equal, nans_equal)] -> (key columns, [lists per agg])
    with lists as [(values, valid)] per group."""
    kcols = [ca.Column.from_numpy(np.asarray(v), valid, t, device="cuda") for t, v, valid in keys]
    t, v, valid = vals
    col = ca.Column.from_numpy(np.asarray(v), valid, t, device="cuda")
    reqs = [ca.GroupByRequest(col, [_agg(*a) for a in aggs] + list(extra))]
    gb = ca.GroupBy(ca.Table(kcols), ca.NullPolicy.INCLUDE if include_null_keys else ca.NullPolicy.EXCLUDE)
    out_keys, results = gb.aggregate(reqs)
    ks = [c.to_numpy() for c in out_keys.columns]
    per_agg = []
    for c in results[0].columns[: len(aggs)]:
        assert isinstance(c, ca.ListColumn) and c.type_id == ca.TypeId.LIST
        assert c.child.type_id == t, "list elements keep the value type"
        off, (cv, cvalid) = c.to_numpy()
        assert off.size == c.size + 1 and off[0] == 0 and np.all(np.diff(off) >= 0)
        assert off[-1] == c.child.size
        cvalid = np.ones(cv.size, bool) if cvalid is None else np.asarray(cvalid, bool)
        assert c.child.null_count == int((~cvalid).sum())
        per_agg.append([(cv[off[g]:off[g + 1]], cvalid[off[g]:off[g + 1]]) for g in range(c.size)])
    return ks, per_agg, results[0].columns[len(aggs):]


@pytest.mark.parametrize("case", COLLECT_CASES, ids=[c["id"] for c in COLLECT_CASES])
def test_reference_collect_fixture_on_gpu(gpu, case):
    ks, lists, _ = gpu_collect([case["keys"]], case["vals"], [case["agg"]])
    G.check_collect_case(case, ks[0], lists[0], sort_sets=False)


def _compare(keys, vals, aggs, include_null_keys=False):
    gk, glists, _ = gpu_collect(keys, vals, aggs, include_null_keys)
    kcols = [O.HostColumn(t, v, valid) for t, v, valid in keys]
    vcol = O.HostColumn(*vals)
    for a, gl in zip(aggs, glists):
        ok, ol = O.collect(kcols, vcol, *a, include_null_keys=include_null_keys)
        assert len(gl) == len(ol), f"{a}: {len(gl)} groups vs {len(ol)}"
        gorder = O.sort_order(gk)
        oorder = O.sort_order(ok)
        for gi, oi in zip(gorder, oorder):
            want = ol[oi]
            if a[0] == O.COLLECT_SET:
                want = O.sort_list(*want)
            assert G.same_list(gl[gi], want), f"{a}: group {gi}: {gl[gi]} vs {want}"


def _values(rng, vt, n, ties=True):
    if vt in (O.FLOAT32, O.FLOAT64):
        pool = np.array([-np.inf, -1.5, -0.0, 0.0, 2.5, 1e30, np.inf, np.nan], dtype=O.NP[vt])
        if ties:
            return pool[rng.integers(0, len(pool), n)]
        return (rng.standard_normal(n) * 50).astype(O.NP[vt])
    info = np.iinfo(O.NP[vt])
    if ties:
        return rng.integers(max(info.min, -6), min(info.max, 6), n, endpoint=True).astype(O.NP[vt])
    return rng.integers(info.min, info.max, n, endpoint=True, dtype=O.NP[vt])


ALL_AGGS = [(O.COLLECT_LIST, 1, 0, 0), (O.COLLECT_LIST, 0, 0, 0), (O.COLLECT_SET, 1, 0, 0), (O.COLLECT_SET, 1, 1, 0),
            (O.COLLECT_SET, 0, 0, 0), (O.COLLECT_SET, 1, 0, 1), (O.COLLECT_SET, 1, 1, 1)]


@pytest.mark.parametrize("vt", [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT8, O.UINT16, O.UINT32, O.UINT64, O.FLOAT32,
                                O.FLOAT64, 12, 16, 21])
def test_value_types(gpu, vt):
    rng = np.random.default_rng(1000 + vt)
    n = 40_000
    keys = rng.integers(0, 2500, n).astype(np.int32)
    for ties in (True, False):
        _compare([(O.INT32, keys, None)], (vt, _values(rng, vt, n, ties), rng.random(n) > 0.1), ALL_AGGS)


def test_signed_zero_and_nan_payloads(gpu):
    """A set keeps one of -0.0 / +0.0 (equal under cudf's comparator; this
    library keeps -0.0 when the group has it) and one NaN or every NaN
    whatever its sign and payload."""
    rng = np.random.default_rng(1010)
    n = 20_000
    keys = rng.integers(0, 300, n).astype(np.int64)
    nans = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0x7FC0BEEF], dtype=np.uint32).view(np.float32)
    pool = np.concatenate([nans, np.array([-0.0, 0.0, 1.0], np.float32)])
    v = pool[rng.integers(0, len(pool), n)]
    _compare([(O.INT64, keys, None)], (O.FLOAT32, v, None), ALL_AGGS)
    _, lists, _ = gpu_collect([(O.INT64, keys, None)], (O.FLOAT32, v, None), [(O.COLLECT_SET, 1, 0, 0)])
    for vals, _valid in lists[0]:
        zeros = vals[vals == 0]
        assert zeros.size <= 1
        assert np.isnan(vals).sum() <= 1


PATHS = ["auto", "partitioned", "direct", "global"]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("card", [5, 30_000])
def test_every_execution(gpu, monkeypatch, path, card):
    if path != "auto":
        monkeypatch.setenv("MGB_GROUPBY_PATH", path)
    rng = np.random.default_rng(1020 + card)
    n = 200_000
    keys = rng.integers(-2 ** 40, 2 ** 40, card)[rng.integers(0, card, n)]
    _compare([(O.INT64, keys, None)], (O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.05), ALL_AGGS)


def test_composite_and_null_keys(gpu):
    rng = np.random.default_rng(1030)
    n = 60_000
    a = rng.integers(0, 60, n).astype(np.int32)
    b = rng.integers(0, 10, n).astype(np.float64) / 2
    keys = [(O.INT32, a, rng.random(n) > 0.05), (O.FLOAT64, b, None)]
    vals = (O.INT64, _values(rng, O.INT64, n), rng.random(n) > 0.1)
    _compare(keys, vals, ALL_AGGS)
    _compare(keys, vals, ALL_AGGS, include_null_keys=True)


def test_empty_and_all_null(gpu):
    e = np.zeros(0, np.int32)
    ks, lists, _ = gpu_collect([(O.INT32, e, None)], (O.FLOAT64, np.zeros(0), None), ALL_AGGS)
    assert ks[0][0].size == 0 and all(len(x) == 0 for x in lists)
    # every key null and excluded: no groups
    k = np.arange(100, dtype=np.int32)
    ks, lists, _ = gpu_collect([(O.INT32, k, np.zeros(100, bool))], (O.INT32, k, None), ALL_AGGS)
    assert ks[0][0].size == 0 and all(len(x) == 0 for x in lists)
    # every value null
    keys = (np.arange(1000) % 7).astype(np.int32)
    _compare([(O.INT32, keys, None)], (O.INT16, np.zeros(1000, np.int16), np.zeros(1000, bool)), ALL_AGGS)


def test_beside_hash_and_sort_aggregations(gpu):
    rng = np.random.default_rng(1040)
    n = 100_000
    keys = rng.integers(0, 9000, n).astype(np.int32)
    v = _values(rng, O.FLOAT64, n, ties=False)
    valid = rng.random(n) > 0.1
    extra = [ca.aggregation.sum(), ca.aggregation.median(), ca.aggregation.nunique(), ca.aggregation.count()]
    aggs = [(O.COLLECT_LIST, 1, 0, 0), (O.COLLECT_SET, 0, 0, 0)]
    gk, glists, others = gpu_collect([(O.INT32, keys, None)], (O.FLOAT64, v, valid), aggs, extra=extra)
    ok, ol = O.collect([O.HostColumn(O.INT32, keys)], O.HostColumn(O.FLOAT64, v, valid), O.COLLECT_LIST, 1)
    g_of = {k: i for i, k in enumerate(gk[0][0].tolist())}
    for k, lst in zip(ok[0][0].tolist(), ol):
        assert G.same_list(glists[0][g_of[k]], lst)
    sums, _ = others[0].to_numpy()
    counts, _ = others[3].to_numpy()
    for k, (lv, lval) in zip(ok[0][0].tolist(), ol):
        assert counts[g_of[k]] == int(lval.sum())
        assert np.isclose(sums[g_of[k]], lv[lval].sum(), rtol=1e-12, atol=1e-9)


def test_two_million_rows(gpu):
    rng = np.random.default_rng(1050)
    n = 2_000_000
    keys = rng.integers(0, 150_000, n).astype(np.int64) * 17
    v = rng.integers(0, 40, n).astype(np.int32)
    _compare([(O.INT64, keys, None)], (O.INT32, v, rng.random(n) > 0.05),
             [(O.COLLECT_LIST, 1, 0, 0), (O.COLLECT_SET, 1, 0, 0)])
