"""GPU parity: the HIP path through the C ABI against the reference's literal
fixtures and against the oracle on seeded random inputs."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as G
from tests.helpers import abs_sums, assert_same, gpu_groupby, oracle_groupby, sorted_result

pytestmark = pytest.mark.gpu

CASES = G.load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["id"] for c in CASES])
def test_reference_fixture(gpu, case):
    kt, kv, kvalid = case["keys"]
    keys, res = gpu_groupby([(kt, kv, kvalid)], [(case["vals"], [case["kind"]])],
                            include_null_keys=case["include_null_keys"])
    G.check_case(case, keys, res)


def _random_case(rng, n, key_type, card, val_types, null_frac, kinds, key_null_frac=0.0):
    if key_type in (O.FLOAT32, O.FLOAT64):
        pool = rng.standard_normal(card).astype(O.NP[key_type])
    else:
        info = np.iinfo(O.NP[key_type])
        pool = rng.integers(max(info.min, -2 ** 62), min(info.max, 2 ** 62), card, dtype=np.int64,
                            endpoint=True).astype(O.NP[key_type])
    keys = pool[rng.integers(0, card, n)]
    kvalid = (rng.random(n) >= key_null_frac) if key_null_frac > 0 else None
    reqs = []
    for vt in val_types:
        if vt in (O.FLOAT32, O.FLOAT64):
            v = rng.standard_normal(n).astype(O.NP[vt])
        elif vt == O.BOOL8:
            v = rng.integers(0, 2, n).astype(np.uint8)
        else:
            info = np.iinfo(O.NP[vt])
            lo, hi = max(info.min, -1_000_000), min(info.max, 1_000_000)
            v = rng.integers(lo, hi, n, endpoint=True).astype(O.NP[vt])
        valid = (rng.random(n) >= null_frac) if null_frac > 0 else None
        reqs.append(((vt, v, valid), kinds))
    return [(key_type, keys, kvalid)], reqs


PATHS = ["auto", "global", "direct", "partitioned", "partitioned-general"]


@pytest.fixture(params=PATHS)
def path(request, monkeypatch):
    """Runs a test once per execution of the groupby (the library reads
    MGB_GROUPBY_PATH on every call; "auto" picks by the cardinality estimate;
    "partitioned-general" turns the compact records off, MGB_LEAN=off)."""
    monkeypatch.delenv("MGB_LEAN", raising=False)
    if request.param == "partitioned-general":
        monkeypatch.setenv("MGB_GROUPBY_PATH", "partitioned")
        monkeypatch.setenv("MGB_LEAN", "off")
    elif request.param != "auto":
        monkeypatch.setenv("MGB_GROUPBY_PATH", request.param)
    else:
        monkeypatch.delenv("MGB_GROUPBY_PATH", raising=False)
    return request.param


def _compare(keys, reqs, include_null_keys=False):
    gk, gr = gpu_groupby(keys, reqs, include_null_keys)
    ok, orr = oracle_groupby(keys, reqs, include_null_keys)
    gk, gr = sorted_result(gk, gr)
    ok, orr = sorted_result(ok, orr)
    sums = abs_sums(keys, reqs, include_null_keys)
    assert len(gk[0][0]) == len(ok[0][0]), "group count differs"
    for c, ((gkv, gkvalid), (okv, okvalid)) in enumerate(zip(gk, ok)):
        assert_same((gkv, gkvalid), (okv, okvalid), None, what=f"key column {c}")
    for r, (gblk, oblk) in enumerate(zip(gr, orr)):
        sk, s = sums[r]
        s_sorted = s[O.sort_order(sk)]
        # float32 sources (their SUM and the SUM inside MEAN are float32 in cudf)
        # are held to the float32 bar; everything else to its result type's bar
        rtol = 1e-5 if reqs[r][0][0] == O.FLOAT32 else None
        for a, (g, o) in enumerate(zip(gblk, oblk)):
            assert_same(g, o, None, abs_sum=s_sorted, rtol=rtol, what=f"request {r} agg {reqs[r][1][a]}")


ALL_KINDS = [O.SUM, O.MIN, O.MAX, O.COUNT_VALID, O.COUNT_ALL, O.MEAN, O.SUM_OF_SQUARES]


@pytest.mark.parametrize("key_type", [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT32, O.UINT64, O.FLOAT32, O.FLOAT64])
def test_key_types(gpu, path, key_type):
    rng = np.random.default_rng(100 + key_type)
    card = 100 if key_type in (O.INT8,) else 5000
    keys, reqs = _random_case(rng, 50_000, key_type, card, [O.INT32, O.FLOAT64], 0.1, ALL_KINDS)
    _compare(keys, reqs)


@pytest.mark.parametrize("val_type", [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT8, O.UINT16, O.UINT32, O.UINT64,
                                      O.FLOAT32, O.FLOAT64, O.BOOL8])
def test_value_types(gpu, path, val_type):
    rng = np.random.default_rng(200 + val_type)
    kinds = [O.SUM, O.MIN, O.MAX, O.COUNT_VALID, O.COUNT_ALL, O.MEAN]
    keys, reqs = _random_case(rng, 40_000, O.INT32, 3000, [val_type], 0.2, kinds)
    _compare(keys, reqs)


@pytest.mark.parametrize("card", [1, 16, 1000, 100_000])
def test_cardinality(gpu, path, card):
    rng = np.random.default_rng(300 + card)
    keys, reqs = _random_case(rng, 200_000, O.INT64, card, [O.INT32, O.FLOAT64, O.FLOAT32], 0.05,
                              [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN])
    _compare(keys, reqs)


def test_null_keys_exclude_and_include(gpu, path):
    rng = np.random.default_rng(11)
    keys, reqs = _random_case(rng, 30_000, O.INT32, 500, [O.INT64], 0.1, ALL_KINDS, key_null_frac=0.1)
    _compare(keys, reqs, include_null_keys=False)
    _compare(keys, reqs, include_null_keys=True)


def test_composite_keys(gpu, path):
    rng = np.random.default_rng(12)
    n = 100_000
    a = rng.integers(0, 300, n).astype(np.int32)
    b = rng.integers(0, 50, n).astype(np.int64) * 1_000_000_007
    c = rng.integers(0, 3, n).astype(np.float64)
    v = rng.standard_normal(n)
    keys = [(O.INT32, a, None), (O.INT64, b, None), (O.FLOAT64, c, rng.random(n) > 0.05)]
    reqs = [((O.FLOAT64, v, rng.random(n) > 0.1), [O.SUM, O.COUNT_VALID, O.MAX, O.MEAN])]
    _compare(keys, reqs)
    _compare(keys, reqs, include_null_keys=True)


def test_sentinel_key_values(gpu, path):
    # INT64_MIN is the empty-slot header of single integer keys; it must still group.
    n = 10_000
    rng = np.random.default_rng(13)
    keys = rng.choice(np.array([np.iinfo(np.int64).min, -1, 0, 1, np.iinfo(np.int64).max], dtype=np.int64), n)
    v = rng.integers(-5, 5, n).astype(np.int32)
    _compare([(O.INT64, keys, None)], [((O.INT32, v, None), [O.SUM, O.COUNT_VALID, O.MIN, O.MAX])])
    ukeys = keys.view(np.uint64)
    _compare([(O.UINT64, ukeys, None)], [((O.INT32, v, None), [O.SUM, O.COUNT_VALID])])


def test_float_key_normalisation(gpu):
    # -0.0 == +0.0 and NaN == NaN as keys (row_operators.cuh:1285-1315)
    keys = np.array([0.0, -0.0, np.nan, -np.nan, 1.5, 1.5, np.nan], dtype=np.float64)
    v = np.arange(7, dtype=np.int64)
    gk, gr = gpu_groupby([(O.FLOAT64, keys, None)], [((O.INT64, v, None), [O.SUM, O.COUNT_VALID])])
    assert len(gk[0][0]) == 3
    got = {("nan" if np.isnan(k) else float(k)): (int(s), int(c))
           for k, s, c in zip(gk[0][0], gr[0][0][0], gr[0][1][0])}
    assert got == {0.0: (1, 2), "nan": (2 + 3 + 6, 3), 1.5: (9, 2)}


def test_duplicate_requests(gpu):
    # the same column in two requests and the same agg twice (keys_tests.cpp:367-421)
    keys = np.array([1, 2, 3, 1, 2, 2, 1, 3, 3, 2], dtype=np.int32)
    vals = np.arange(10, dtype=np.int32)
    gk, gr = gpu_groupby([(O.INT32, keys, None)],
                         [((O.INT32, vals, None), [O.SUM, O.SUM]), ((O.INT32, vals, None), [O.SUM, O.MEAN])])
    order = np.argsort(gk[0][0])
    for blk in gr:
        assert list(blk[0][0][order]) == [9, 19, 17]
    assert list(gr[0][1][0][order]) == [9, 19, 17]
    assert np.allclose(gr[1][1][0][order], [3.0, 4.75, 17 / 3])


def test_nan_values_min_max_terminate(gpu):
    # min_tests.cpp:524-547 / max_tests.cpp:527-551: NaN values must not hang
    for t in (O.FLOAT32, O.FLOAT64):
        keys = np.array([1, 1], dtype=np.int32)
        v = np.array([np.nan, np.nan], dtype=O.NP[t])
        gk, gr = gpu_groupby([(O.INT32, keys, None)], [((t, v, None), [O.MIN, O.MAX])])
        assert len(gk[0][0]) == 1
        assert np.isnan(gr[0][0][0][0]) and np.isnan(gr[0][1][0][0])


def test_nan_skipped_when_other_values(gpu):
    # float MIN/MAX ignore NaN unless every valid value of the group is NaN (pandas skipna)
    keys = np.array([1, 1, 1, 2, 2], dtype=np.int32)
    v = np.array([np.nan, 3.0, -1.0, np.nan, np.nan], dtype=np.float64)
    gk, gr = gpu_groupby([(O.INT32, keys, None)], [((O.FLOAT64, v, None), [O.MIN, O.MAX])])
    order = np.argsort(gk[0][0])
    mn, mx = gr[0][0][0][order], gr[0][1][0][order]
    assert mn[0] == -1.0 and mx[0] == 3.0
    assert np.isnan(mn[1]) and np.isnan(mx[1])


def test_c2_shape_slice(gpu, path):
    """2M rows of the 100M-row benchmark's distribution (int64 keys over 200K
    distinct values, int32/f64/f32 values 5% null, 5 aggs each)."""
    rng = np.random.default_rng(42)
    n = 2_000_000
    uniq = np.unique(rng.integers(-2 ** 63, 2 ** 63 - 1, 200_000, dtype=np.int64))
    keys = uniq[rng.integers(0, uniq.size, n)]
    kinds = [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN]
    reqs = [((O.INT32, rng.integers(-1_000_000, 1_000_001, n).astype(np.int32), rng.random(n) >= 0.05), kinds),
            ((O.FLOAT64, rng.standard_normal(n), rng.random(n) >= 0.05), kinds),
            ((O.FLOAT32, rng.standard_normal(n).astype(np.float32), rng.random(n) >= 0.05), kinds)]
    _compare([(O.INT64, keys, None)], reqs)


def _zipf_ranks(rng, n, n_keys, s):
    u = rng.random(n)
    return np.clip(np.floor((1.0 + u * (n_keys ** (1 - s) - 1.0)) ** (1.0 / (1 - s))).astype(np.int64) - 1,
                   0, n_keys - 1)


def test_zipf_composite_keys(gpu, path):
    """The shape-stress workload at 2M rows: composite (int32, int64) keys,
    Zipf s=1.1 over 500K tuples (the hottest key holds ~12% of the rows, so
    its partition is split over several work items), f64 sum/count/max."""
    rng = np.random.default_rng(31)
    n, n_keys = 2_000_000, 500_000
    ta = rng.integers(-2 ** 31, 2 ** 31 - 1, n_keys).astype(np.int32)
    tb = rng.integers(-2 ** 62, 2 ** 62, n_keys).astype(np.int64)
    perm = rng.permutation(n_keys)
    idx = perm[_zipf_ranks(rng, n, n_keys, 1.1)]
    keys = [(O.INT32, ta[idx], None), (O.INT64, tb[idx], None)]
    reqs = [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX])]
    _compare(keys, reqs)


def test_hot_key_partition_split(gpu, path):
    """One int64 key holds 40% of 1M rows: far more than one work item, so
    the partition is aggregated by several workgroups whose partial groups
    meet in the global table."""
    rng = np.random.default_rng(32)
    n = 1_000_000
    keys = rng.integers(0, 50_000, n).astype(np.int64)
    keys[rng.random(n) < 0.4] = 123456789
    reqs = [((O.INT64, rng.integers(-1000, 1000, n).astype(np.int64), rng.random(n) > 0.1),
             [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN]),
            ((O.FLOAT32, rng.standard_normal(n).astype(np.float32), None), [O.SUM, O.MIN, O.MAX, O.COUNT_ALL])]
    _compare([(O.INT64, keys, None)], reqs)


def test_few_partitions_many_rows_split(gpu, path):
    """3M rows over 2,500 groups: partitions are kept at 64 groups or more,
    so each of about 94K rows is cut into two work items, whose partial
    groups meet in the global table."""
    rng = np.random.default_rng(34)
    n = 3_000_000
    pool = rng.integers(-2 ** 62, 2 ** 62, 2500, dtype=np.int64)
    keys = pool[rng.integers(0, 2500, n)]
    reqs = [((O.FLOAT64, rng.standard_normal(n), rng.random(n) > 0.02), [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN]),
            ((O.INT32, rng.integers(-1_000_000, 1_000_000, n).astype(np.int32), rng.random(n) > 0.02),
             [O.SUM, O.MIN, O.MAX, O.MEAN])]
    _compare([(O.INT64, keys, None)], reqs)


def test_low_cardinality_16_keys(gpu, path):
    """The LDS-path workload at 4M rows: int32 key with 16 distinct values."""
    rng = np.random.default_rng(33)
    n = 4_000_000
    uniq = rng.integers(-2 ** 31, 2 ** 31 - 1, 16).astype(np.int32)
    keys = uniq[rng.integers(0, 16, n)]
    reqs = [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX])]
    _compare([(O.INT32, keys, None)], reqs)


@pytest.mark.parametrize("scale", ["0.001", "0.05"])
def test_underestimate_retries(gpu, path, scale, monkeypatch):
    """The estimate scaled far below the true group count: the direct and
    global executions' tables overflow and are rebuilt larger, the
    partitioned execution's spill table overflows and the rows are
    repartitioned finer (falling back to the global execution after four
    attempts). Every attempt ends (bounded probing, overflow flag) and the
    result is the oracle's."""
    monkeypatch.setenv("MGB_ESTIMATE_SCALE", scale)
    rng = np.random.default_rng(35)
    keys, reqs = _random_case(rng, 400_000, O.INT64, 100_000, [O.INT32, O.FLOAT64], 0.05,
                              [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN])
    _compare(keys, reqs)


def test_underestimate_retries_zipf_hot_key(gpu, path, monkeypatch):
    """The composite-key Zipf shape with a 1000x underestimate: the hot key's
    partition is split and the spill table overflows on the first attempts."""
    monkeypatch.setenv("MGB_ESTIMATE_SCALE", "0.001")
    rng = np.random.default_rng(36)
    n, n_keys = 1_000_000, 200_000
    ta = rng.integers(-2 ** 31, 2 ** 31 - 1, n_keys).astype(np.int32)
    tb = rng.integers(-2 ** 62, 2 ** 62, n_keys).astype(np.int64)
    idx = rng.permutation(n_keys)[_zipf_ranks(rng, n, n_keys, 1.1)]
    _compare([(O.INT32, ta[idx], None), (O.INT64, tb[idx], None)],
             [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX])])


def test_product_against_oracle(gpu, path):
    """PRODUCT of integer and float columns (integer products wrap modulo
    2^64 in any order, so they are bit-exact; float32 products are held at
    float32 precision in every fold)."""
    rng = np.random.default_rng(37)
    n, card = 400_000, 2000
    pool = rng.integers(-2 ** 62, 2 ** 62, card, dtype=np.int64)
    keys = pool[rng.integers(0, card, n)]
    reqs = [((O.INT8, rng.integers(-3, 4, n).astype(np.int8), rng.random(n) > 0.1), [O.PRODUCT, O.SUM]),
            ((O.INT32, rng.integers(-1_000_000, 1_000_000, n).astype(np.int32), rng.random(n) > 0.05), [O.PRODUCT]),
            ((O.INT64, rng.integers(-2 ** 40, 2 ** 40, n, dtype=np.int64), None), [O.PRODUCT, O.MAX]),
            ((O.UINT16, rng.integers(0, 2 ** 16, n).astype(np.uint16), rng.random(n) > 0.05), [O.PRODUCT]),
            ((O.FLOAT32, rng.uniform(0.95, 1.05, n).astype(np.float32), rng.random(n) > 0.05), [O.PRODUCT]),
            ((O.FLOAT64, rng.uniform(0.9, 1.1, n), rng.random(n) > 0.05), [O.PRODUCT, O.SUM])]
    _compare([(O.INT64, keys, None)], reqs)


def test_product_hot_key_split(gpu, path):
    """PRODUCT where one key holds 40% of 1M rows: its partition is split
    over work items whose partial products meet in the global table. Float32
    factors are +-1 with rare 2 and 0.5, so every order gives the exact
    product."""
    rng = np.random.default_rng(38)
    n = 1_000_000
    keys = rng.integers(0, 30_000, n).astype(np.int64)
    keys[rng.random(n) < 0.4] = 77
    f32 = rng.choice(np.array([1.0, -1.0, 2.0, 0.5], np.float32), n, p=[0.4995, 0.4995, 0.0005, 0.0005])
    reqs = [((O.FLOAT32, f32, None), [O.PRODUCT, O.COUNT_VALID]),
            ((O.INT64, rng.integers(-9, 10, n, dtype=np.int64), rng.random(n) > 0.2), [O.PRODUCT]),
            ((O.FLOAT64, rng.uniform(0.999, 1.001, n), None), [O.PRODUCT])]
    gk, gr = gpu_groupby([(O.INT64, keys, None)], reqs)
    ok, orr = oracle_groupby([(O.INT64, keys, None)], reqs)
    gk, gr = sorted_result(gk, gr)
    ok, orr = sorted_result(ok, orr)
    assert np.array_equal(gk[0][0], ok[0][0])
    assert np.array_equal(gr[0][0][0], orr[0][0][0]), "float32 products of +-1, 2, 0.5 are exact in any order"
    assert np.array_equal(gr[1][0][0], orr[1][0][0]) and np.array_equal(gr[1][0][1], orr[1][0][1])
    assert_same(gr[2][0], orr[2][0], None, what="float64 product")
