"""GPU parity of the compact partitioned execution (an integer key of one or
two 4- or 8-byte columns, up to three value columns of SUM / MIN / MAX /
counts, records of 16 or 32 bytes holding the row hash and validity bits):
every key and value width, two-column keys whose 8-byte parts are shared by
many groups, null values and null keys, the empty-header key, spills into the
global table, split partitions and 8192 partitions, against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_parity import _compare, _random_case, _zipf_ranks

pytestmark = pytest.mark.gpu

KINDS = [O.SUM, O.MIN, O.MAX, O.COUNT_VALID, O.COUNT_ALL, O.MEAN]


@pytest.fixture(autouse=True)
def partitioned(monkeypatch):
    monkeypatch.setenv("MGB_GROUPBY_PATH", "partitioned")
    monkeypatch.delenv("MGB_LEAN", raising=False)
    monkeypatch.delenv("MGB_ESTIMATE_SCALE", raising=False)
    monkeypatch.delenv("MGB_HOT_KEYS", raising=False)


# (key type, value types): records of 16 bytes (4-byte key + hash + 8 bytes
# of values) and of 32 bytes, every value width, three columns
LAYOUTS = [
    (O.INT32, [O.INT32]),
    (O.UINT32, [O.FLOAT32, O.INT16]),
    (O.INT32, [O.INT64]),
    (O.INT32, [O.INT8, O.UINT16, O.BOOL8]),
    (O.INT64, [O.INT32]),
    (O.INT64, [O.FLOAT64]),
    (O.UINT64, [O.UINT8, O.INT16, O.UINT32]),
    (O.INT64, [O.INT32, O.FLOAT64, O.FLOAT32]),
    (O.INT64, [O.FLOAT64, O.INT64]),
    (O.INT32, [O.FLOAT64, O.UINT64, O.INT32]),
]


@pytest.mark.parametrize("layout", LAYOUTS, ids=[f"k{k}-" + "-".join(map(str, v)) for k, v in LAYOUTS])
def test_layouts(gpu, layout):
    key_type, val_types = layout
    rng = np.random.default_rng(500 + key_type * 31 + sum(val_types))
    keys, reqs = _random_case(rng, 300_000, key_type, 40_000, val_types, 0.1, KINDS)
    _compare(keys, reqs)


def test_null_keys_dropped(gpu):
    rng = np.random.default_rng(501)
    keys, reqs = _random_case(rng, 200_000, O.INT64, 30_000, [O.INT32, O.FLOAT64], 0.05, KINDS, key_null_frac=0.1)
    _compare(keys, reqs, include_null_keys=False)


def test_no_nullable_values(gpu):
    rng = np.random.default_rng(502)
    keys, reqs = _random_case(rng, 200_000, O.INT32, 30_000, [O.FLOAT32, O.INT64], 0.0, KINDS)
    _compare(keys, reqs)


def test_only_row_counts(gpu):
    rng = np.random.default_rng(503)
    keys, reqs = _random_case(rng, 200_000, O.INT64, 50_000, [O.INT32], 0.2, [O.COUNT_ALL, O.COUNT_VALID])
    _compare(keys, reqs)


def test_extreme_keys(gpu):
    """INT64_MIN is the empty-slot header of integer keys (its group has a
    record of its own); the other extremes of every key width group as usual."""
    rng = np.random.default_rng(504)
    n = 400_000
    for kt, dt in ((O.INT64, np.int64), (O.UINT64, np.uint64), (O.INT32, np.int32), (O.UINT32, np.uint32)):
        info = np.iinfo(dt)
        base = rng.integers(0, 2 ** 20, 60_000).astype(dt)
        pool = np.concatenate([base, np.array([info.min, info.max, 0, 1], dtype=dt)])
        keys = pool[rng.integers(0, pool.size, n)]
        keys[rng.random(n) < 0.01] = info.min
        v = rng.standard_normal(n)
        _compare([(kt, keys, None)], [((O.FLOAT64, v, rng.random(n) > 0.1), KINDS)])


def test_nan_and_signed_zero_values(gpu):
    rng = np.random.default_rng(505)
    n = 300_000
    keys = rng.integers(0, 40_000, n).astype(np.int64)
    v = rng.standard_normal(n)
    v[rng.random(n) < 0.05] = np.nan
    v[rng.random(n) < 0.05] = -0.0
    v[rng.random(n) < 0.02] = np.inf
    f = v.astype(np.float32)
    _compare([(O.INT64, keys, None)], [((O.FLOAT64, v, rng.random(n) > 0.1), [O.SUM, O.MIN, O.MAX, O.COUNT_VALID]),
                                       ((O.FLOAT32, f, None), [O.MIN, O.MAX, O.MEAN])])


@pytest.mark.parametrize("scale", ["0.05", "0.001"])
def test_underestimate_spills_and_retries(gpu, scale, monkeypatch):
    """A scaled-down estimate makes the items' LDS tables overflow: their
    groups spill into the global table, which itself overflows at the
    smaller scale, and the call is repeated with more partitions."""
    monkeypatch.setenv("MGB_ESTIMATE_SCALE", scale)
    rng = np.random.default_rng(506)
    keys, reqs = _random_case(rng, 500_000, O.INT64, 200_000, [O.INT32, O.FLOAT64, O.FLOAT32], 0.05,
                              [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN])
    _compare(keys, reqs)


def test_zipf_hot_keys_split(gpu):
    """Zipf s=1.1 over 300K int64 keys at 3M rows: the hottest partitions are
    cut into several work items whose groups meet in the global table."""
    rng = np.random.default_rng(507)
    n, n_keys = 3_000_000, 300_000
    pool = rng.integers(-2 ** 62, 2 ** 62, n_keys, dtype=np.int64)
    keys = pool[rng.permutation(n_keys)[_zipf_ranks(rng, n, n_keys, 1.1)]]
    reqs = [((O.INT32, rng.integers(-1000, 1000, n).astype(np.int32), rng.random(n) > 0.05), KINDS),
            ((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.MAX])]
    _compare([(O.INT64, keys, None)], reqs)


def test_many_partitions(gpu):
    """2M groups over 4M rows: 4096 partitions (the most the compact scatter
    stages; its tile shrinks to fit the partition counters)."""
    rng = np.random.default_rng(508)
    keys, reqs = _random_case(rng, 4_000_000, O.INT64, 2_000_000, [O.INT32, O.FLOAT64], 0.05,
                              [O.SUM, O.COUNT_VALID, O.MIN, O.MAX])
    _compare(keys, reqs)


def test_same_result_as_general_records(gpu, monkeypatch):
    """The compact and the general partitioned records give the same groups
    and bit-identical integer results."""
    from tests.helpers import gpu_groupby, sorted_result

    rng = np.random.default_rng(509)
    keys, reqs = _random_case(rng, 1_000_000, O.INT64, 100_000, [O.INT32, O.INT64], 0.05,
                              [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.COUNT_ALL])
    a = sorted_result(*gpu_groupby(keys, reqs))
    monkeypatch.setenv("MGB_LEAN", "off")
    b = sorted_result(*gpu_groupby(keys, reqs))
    assert np.array_equal(a[0][0][0], b[0][0][0])
    for blk_a, blk_b in zip(a[1], b[1]):
        for (va, ma), (vb, mb) in zip(blk_a, blk_b):
            assert np.array_equal(va, vb)
            assert (ma is None and mb is None) or np.array_equal(ma, mb)


# ---- two-column integer keys: part A (the first 8 key bytes) in the LDS
# header, part B and the first record's index in the record's tail

def _pair_keys(rng, n, types, card, shared_a=0, key_null_frac=0.0):
    """Two key columns drawn as `card` distinct tuples; with shared_a > 0 the
    8-byte part of the key takes only that many values, so many groups share
    their header and differ in the tail."""
    cols = []
    for c, t in enumerate(types):
        info = np.iinfo(O.NP[t])
        lo, hi = max(info.min, -2 ** 62), min(info.max, 2 ** 62)
        cols.append(rng.integers(lo, hi, card, dtype=np.int64, endpoint=True).astype(O.NP[t]))
    if shared_a:
        wide = 0 if O.NP[types[0]]().itemsize >= O.NP[types[1]]().itemsize else 1
        cols[wide] = cols[wide][rng.integers(0, shared_a, card)]
    idx = rng.integers(0, card, n)
    out = []
    for c, t in enumerate(types):
        valid = (rng.random(n) >= key_null_frac) if key_null_frac > 0 else None
        out.append((t, cols[c][idx], valid))
    return out


PAIR_LAYOUTS = [
    ((O.INT32, O.INT64), [O.FLOAT64]),
    ((O.INT64, O.INT32), [O.INT32, O.FLOAT32]),
    ((O.INT64, O.INT64), [O.INT32]),
    ((O.INT32, O.INT32), [O.FLOAT64, O.INT64]),
    ((O.UINT32, O.UINT64), [O.INT32, O.INT16, O.UINT32]),
    ((O.UINT64, O.INT64), []),
    ((O.INT32, O.UINT32), [O.INT8, O.UINT16, O.FLOAT32]),
]


@pytest.mark.parametrize("layout", PAIR_LAYOUTS,
                         ids=[f"k{a}.{b}-" + "-".join(map(str, v)) for (a, b), v in PAIR_LAYOUTS])
@pytest.mark.parametrize("shared_a", [0, 300])
def test_two_column_layouts(gpu, layout, shared_a):
    key_types, val_types = layout
    rng = np.random.default_rng(600 + 7 * sum(key_types) + sum(val_types) + shared_a)
    n = 300_000
    keys = _pair_keys(rng, n, key_types, 60_000, shared_a)
    _, reqs = _random_case(rng, n, O.INT32, 10, val_types or [O.INT32], 0.1, KINDS)
    if not val_types:
        reqs = [(reqs[0][0], [O.COUNT_ALL])]
    _compare(keys, reqs)


def test_two_column_null_keys_dropped(gpu):
    rng = np.random.default_rng(610)
    n = 300_000
    keys = _pair_keys(rng, n, (O.INT32, O.INT64), 50_000, key_null_frac=0.05)
    _, reqs = _random_case(rng, n, O.INT32, 10, [O.FLOAT64, O.INT32], 0.05, KINDS)
    _compare(keys, reqs, include_null_keys=False)


def test_two_column_empty_header_part(gpu):
    """Tuples whose 8-byte part equals the empty header (INT64_MIN, and
    (0, INT32_MIN) for two 4-byte columns) take the global table; others
    sharing their second part do not."""
    rng = np.random.default_rng(611)
    n = 200_000
    a = rng.integers(-1000, 1000, n).astype(np.int64)
    b = rng.integers(-50, 50, n).astype(np.int32)
    a[rng.random(n) < 0.05] = np.iinfo(np.int64).min
    v = rng.standard_normal(n)
    _compare([(O.INT64, a, None), (O.INT32, b, None)], [((O.FLOAT64, v, rng.random(n) > 0.1), KINDS)])
    c = rng.integers(-1000, 1000, n).astype(np.int32)
    d = rng.integers(-1000, 1000, n).astype(np.int32)
    hit = rng.random(n) < 0.05
    c[hit] = 0
    d[hit] = np.iinfo(np.int32).min
    _compare([(O.INT32, c, None), (O.INT32, d, None)], [((O.FLOAT64, v, None), KINDS)])


@pytest.mark.parametrize("scale", ["0.05", "0.001"])
def test_two_column_underestimate(gpu, scale, monkeypatch):
    monkeypatch.setenv("MGB_ESTIMATE_SCALE", scale)
    rng = np.random.default_rng(612)
    n = 500_000
    keys = _pair_keys(rng, n, (O.INT32, O.INT64), 200_000, shared_a=5000)
    _, reqs = _random_case(rng, n, O.INT32, 10, [O.FLOAT64, O.INT32], 0.05, [O.SUM, O.COUNT_VALID, O.MAX, O.MIN])
    _compare(keys, reqs)


def test_two_column_zipf(gpu):
    """The shape-stress distribution at 4M rows: (int32, int64) tuples, Zipf
    s=1.1 over 1M of them; hot partitions cut into several items."""
    rng = np.random.default_rng(613)
    n, n_keys = 4_000_000, 1_000_000
    ta = rng.integers(-2 ** 31, 2 ** 31 - 1, n_keys).astype(np.int32)
    tb = rng.integers(-2 ** 62, 2 ** 62, n_keys).astype(np.int64)
    idx = rng.permutation(n_keys)[_zipf_ranks(rng, n, n_keys, 1.1)]
    reqs = [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX])]
    _compare([(O.INT32, ta[idx], None), (O.INT64, tb[idx], None)], reqs)


def test_eight_thousand_partitions(gpu):
    """Enough groups for 8192 partitions (the most the compact scatter
    takes), with one key column and with two."""
    rng = np.random.default_rng(614)
    keys, reqs = _random_case(rng, 7_000_000, O.INT64, 5_500_000, [O.INT32], 0.05, [O.SUM, O.COUNT_VALID, O.MAX])
    _compare(keys, reqs)
    n = 12_000_000
    keys = _pair_keys(rng, n, (O.INT32, O.INT64), 11_000_000)
    _compare(keys, [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX])])


def test_two_column_same_result_as_general_records(gpu, monkeypatch):
    from tests.helpers import gpu_groupby, sorted_result

    rng = np.random.default_rng(615)
    n = 1_000_000
    keys = _pair_keys(rng, n, (O.INT64, O.INT32), 150_000, shared_a=2000)
    _, reqs = _random_case(rng, n, O.INT32, 10, [O.INT32, O.INT64], 0.05,
                           [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.COUNT_ALL])
    a = sorted_result(*gpu_groupby(keys, reqs))
    monkeypatch.setenv("MGB_LEAN", "off")
    b = sorted_result(*gpu_groupby(keys, reqs))
    for (ka, _), (kb, _) in zip(a[0], b[0]):
        assert np.array_equal(ka, kb)
    for blk_a, blk_b in zip(a[1], b[1]):
        for (va, ma), (vb, mb) in zip(blk_a, blk_b):
            assert np.array_equal(va, vb)
            assert (ma is None and mb is None) or np.array_equal(ma, mb)


# ---- hot keys: keys seen 8 times or more in the estimate's sample skip the
# partitioned records and are aggregated by the scatter

def _skewed_keys(rng, n, pool, hot_frac, n_hot):
    """Keys from `pool`: hot_frac of the rows on its first n_hot keys (a few
    of them heavier), the rest uniform over the pool."""
    hot = rng.random(n) < hot_frac
    w = 1.0 / np.arange(1, n_hot + 1) ** 0.7
    idx = np.where(hot, rng.choice(n_hot, n, p=w / w.sum()), rng.integers(0, len(pool), n))
    return pool[idx]


@pytest.mark.parametrize("kt", [O.INT64, O.INT32, O.UINT32])
def test_hot_keys_single_column(gpu, kt):
    rng = np.random.default_rng(620 + kt)
    n = 2_000_000
    info = np.iinfo(O.NP[kt])
    pool = rng.integers(max(info.min, -2 ** 62), min(info.max, 2 ** 62), 400_000, dtype=np.int64).astype(O.NP[kt])
    keys = _skewed_keys(rng, n, pool, 0.6, 120)
    v = rng.standard_normal(n)
    v[rng.random(n) < 0.02] = -0.0
    i32 = rng.integers(-1000, 1000, n).astype(np.int32)
    reqs = [((O.FLOAT64, v, rng.random(n) > 0.05), KINDS), ((O.INT32, i32, rng.random(n) > 0.3), KINDS)]
    _compare([(kt, keys, None)], reqs)


@pytest.mark.parametrize("types", [(O.INT32, O.INT64), (O.INT64, O.INT64), (O.INT32, O.INT32)])
def test_hot_keys_two_columns(gpu, types):
    rng = np.random.default_rng(630 + sum(types))
    n = 2_000_000
    card = 300_000
    cols = [rng.integers(np.iinfo(O.NP[t]).min // 2, np.iinfo(O.NP[t]).max // 2, card, dtype=np.int64).astype(O.NP[t])
            for t in types]
    hot = rng.random(n) < 0.5
    w = 1.0 / np.arange(1, 201) ** 0.8
    idx = np.where(hot, rng.choice(200, n, p=w / w.sum()), rng.integers(0, card, n))
    keys = [(types[0], cols[0][idx], None), (types[1], cols[1][idx], None)]
    reqs = [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX, O.MIN]),
            ((O.INT64, rng.integers(-10 ** 9, 10 ** 9, n), rng.random(n) > 0.1), [O.SUM, O.MIN, O.COUNT_ALL])]
    _compare(keys, reqs)


def test_hot_keys_more_than_the_cap_and_null_keys(gpu):
    """Half the rows on 3000 keys (about 2000 of them seen 8 times or more in
    the estimate's sample): more hot candidates than the 1024 taken (the rest
    are partitioned as usual); null keys dropped."""
    rng = np.random.default_rng(640)
    n = 3_000_000
    pool = rng.integers(-2 ** 40, 2 ** 40, 500_000)
    keys = _skewed_keys(rng, n, pool, 0.5, 3000)
    kvalid = rng.random(n) > 0.05
    _, reqs = _random_case(rng, n, O.INT64, 10, [O.INT32, O.FLOAT32], 0.05, KINDS)
    _compare([(O.INT64, keys, kvalid)], reqs)


def test_hot_keys_same_result_as_without(gpu, monkeypatch):
    from tests.helpers import gpu_groupby, sorted_result

    rng = np.random.default_rng(650)
    n = 2_000_000
    pool = rng.integers(-2 ** 50, 2 ** 50, 300_000)
    keys = [(O.INT64, _skewed_keys(rng, n, pool, 0.7, 64), None)]
    _, reqs = _random_case(rng, n, O.INT32, 10, [O.INT32, O.INT64], 0.05,
                           [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.COUNT_ALL])
    a = sorted_result(*gpu_groupby(keys, reqs))
    monkeypatch.setenv("MGB_HOT_KEYS", "off")
    b = sorted_result(*gpu_groupby(keys, reqs))
    assert np.array_equal(a[0][0][0], b[0][0][0])
    for blk_a, blk_b in zip(a[1], b[1]):
        for (va, ma), (vb, mb) in zip(blk_a, blk_b):
            assert np.array_equal(va, vb)
            assert (ma is None and mb is None) or np.array_equal(ma, mb)
