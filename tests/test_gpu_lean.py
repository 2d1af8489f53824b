"""GPU parity of the compact partitioned execution (a single 4- or 8-byte
integer key, up to three value columns of SUM / MIN / MAX / counts, records
of 16 or 32 bytes holding the row hash and validity bits): every key and
value width, null values and null keys, the empty-header key, spills into the
global table and split partitions, against the oracle."""
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
