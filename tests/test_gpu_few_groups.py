"""The few-groups direct execution (single integer key, every group among
the sample's distinct keys, at most 16 of them, one or two value columns of
SUM / MIN / MAX / counts): lane-striped record copies, with and without the
in-wave collapse of each group's lanes, and turned off, against the oracle;
keys the sample missed take the LDS table's probe."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.test_gpu_parity import _compare

pytestmark = pytest.mark.gpu


@pytest.fixture(params=["on", "collapse", "off"])
def few(request, monkeypatch):
    if request.param == "on":
        monkeypatch.delenv("MGB_FEW_GROUPS", raising=False)
    else:
        monkeypatch.setenv("MGB_FEW_GROUPS", request.param)
    monkeypatch.delenv("MGB_GROUPBY_PATH", raising=False)
    return request.param


def _vals(rng, t, n):
    if t in (O.FLOAT32, O.FLOAT64):
        v = rng.standard_normal(n).astype(O.NP[t])
        special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan], O.NP[t])
        pick = rng.random(n) < 0.01
        v[pick] = special[rng.integers(0, special.size, int(pick.sum()))]
        return v
    if t == O.BOOL8:
        return rng.integers(0, 2, n).astype(np.uint8)
    info = np.iinfo(O.NP[t])
    return rng.integers(max(info.min, -1_000_000), min(info.max, 1_000_000), n, endpoint=True).astype(O.NP[t])


KINDS = [O.SUM, O.COUNT_VALID, O.COUNT_ALL, O.MIN, O.MAX, O.MEAN]


@pytest.mark.parametrize("groups", [1, 5, 16])
@pytest.mark.parametrize("key_type", [O.INT32, O.INT64, O.UINT8, O.INT16])
def test_few_groups_key_types(gpu, few, groups, key_type):
    rng = np.random.default_rng(500 + groups + 10 * key_type)
    n = 3_000_000
    info = np.iinfo(O.NP[key_type])
    pool = np.unique(rng.integers(info.min, info.max, 64, dtype=np.int64, endpoint=True))[:groups]
    keys = pool[rng.integers(0, pool.size, n)].astype(O.NP[key_type])
    reqs = [((O.FLOAT64, _vals(rng, O.FLOAT64, n), rng.random(n) > 0.05), KINDS),
            ((O.INT32, _vals(rng, O.INT32, n), None), [O.SUM, O.MIN, O.MAX, O.COUNT_VALID])]
    _compare([(key_type, keys, None)], reqs)


@pytest.mark.parametrize("val_type", [O.INT8, O.INT16, O.INT64, O.UINT8, O.UINT32, O.UINT64, O.FLOAT32, O.BOOL8])
def test_few_groups_value_types(gpu, few, val_type):
    rng = np.random.default_rng(600 + val_type)
    n = 1_000_000
    keys = rng.integers(-8, 8, n).astype(np.int32)
    kinds = [O.SUM, O.MIN, O.MAX, O.COUNT_VALID, O.MEAN]
    reqs = [((val_type, _vals(rng, val_type, n), rng.random(n) > 0.1), kinds)]
    _compare([(O.INT32, keys, None)], reqs)


def test_few_groups_with_rare_key_missed_by_the_sample(gpu, few):
    """16 common keys plus 40 rows of a 17th key and the key equal to the
    empty header (INT64_MIN): the sample sees 16 keys, the rare rows take the
    table's probe and still meet their own groups."""
    rng = np.random.default_rng(601)
    n = 2_000_000
    pool = rng.integers(-2 ** 62, 2 ** 62, 16, dtype=np.int64)
    keys = pool[rng.integers(0, 16, n)]
    rare = rng.choice(n, 40, replace=False)
    keys[rare[:30]] = 123456789
    keys[rare[30:]] = np.iinfo(np.int64).min
    reqs = [((O.FLOAT64, rng.standard_normal(n), rng.random(n) > 0.02), [O.SUM, O.COUNT_VALID, O.MAX, O.MIN]),
            ((O.INT64, rng.integers(-10 ** 9, 10 ** 9, n, dtype=np.int64), None), [O.SUM, O.MAX])]
    _compare([(O.INT64, keys, None)], reqs)


def test_few_groups_null_keys_dropped(gpu, few):
    rng = np.random.default_rng(602)
    n = 1_000_000
    keys = rng.integers(0, 10, n).astype(np.int32)
    kvalid = rng.random(n) > 0.1
    reqs = [((O.FLOAT32, rng.standard_normal(n).astype(np.float32), rng.random(n) > 0.3), [O.SUM, O.MAX, O.COUNT_ALL])]
    _compare([(O.INT32, keys, kvalid)], reqs)


def test_few_groups_not_taken_for_other_ops(gpu, few):
    """SUM_OF_SQUARES and three value columns leave the few-groups execution
    (the generic direct execution runs them)."""
    rng = np.random.default_rng(604)
    n = 500_000
    keys = rng.integers(0, 7, n).astype(np.int64)
    reqs = [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM_OF_SQUARES, O.SUM]),
            ((O.INT32, rng.integers(-100, 100, n).astype(np.int32), None), [O.MAX]),
            ((O.FLOAT32, rng.standard_normal(n).astype(np.float32), rng.random(n) > 0.5), [O.MIN])]
    _compare([(O.INT64, keys, None)], reqs)


def test_few_groups_shape_stress_16_keys(gpu, few):
    """The 16-key workload's shape at 8M rows: int32 key over 16 values,
    float64 value, sum / count / max."""
    rng = np.random.default_rng(603)
    n = 8_000_000
    uniq = rng.integers(-2 ** 31, 2 ** 31 - 1, 16).astype(np.int32)
    keys = uniq[rng.integers(0, 16, n)]
    _compare([(O.INT32, keys, None)], [((O.FLOAT64, rng.standard_normal(n), None), [O.SUM, O.COUNT_VALID, O.MAX])])


@pytest.mark.parametrize("key_type", [O.INT32, O.INT64])
@pytest.mark.parametrize("val_type", [O.INT32, O.INT64, O.FLOAT32, O.FLOAT64])
def test_few_groups_compiled_shapes(gpu, few, key_type, val_type):
    """One value column of each type the row loop is compiled for, under 4-
    and 8-byte keys, with nulls, NaN, infinities and signed zeros."""
    rng = np.random.default_rng(700 + 10 * key_type + val_type)
    n = 2_000_000
    pool = np.unique(rng.integers(-2 ** 31, 2 ** 31 - 1, 20)).astype(O.NP[key_type])[:12]
    keys = pool[rng.integers(0, pool.size, n)]
    reqs = [((val_type, _vals(rng, val_type, n), rng.random(n) > 0.03), [O.SUM, O.MIN, O.MAX, O.COUNT_VALID, O.MEAN])]
    _compare([(key_type, keys, None)], reqs)
