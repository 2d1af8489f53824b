"""GPU parity of the exported row hashers (cudf::hashing::murmurhash3_x86_32
and spark_murmurhash3_x86_32, cpp/include/cudf/hashing.hpp:90-134) through
the C ABI: against the oracle's restatement over every fixed-width type,
composite rows, nulls, -0.0 / NaN and several seeds; against the Spark
golden vectors; and against the invariants of
cpp/tests/hashing/murmurhash3_x86_32_test.cpp:31-121 (bool truth values and
null rows hash alike whatever their storage holds)."""
from __future__ import annotations

import json
import os

import numpy as np
import pytest

import cudf_amd as ca
from cudf_amd import hashing
from oracle import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VECTORS = json.load(open(os.path.join(HERE, "golden", "hash_vectors.json")))
TS_DAYS, TS_MS = 12, 14


def _values(rng, t, n):
    if t in (O.FLOAT32, O.FLOAT64):
        v = rng.standard_normal(n).astype(O.NP[t])
        special = np.array([0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf], O.NP[t])
        pick = rng.random(n) < 0.2
        v[pick] = special[rng.integers(0, special.size, int(pick.sum()))]
        # NaNs with other payloads: every NaN hashes as the canonical one
        if t == O.FLOAT64:
            bits = v.view(np.uint64)
            odd = rng.random(n) < 0.02
            bits[odd] = np.uint64(0x7ff0000000000001) | rng.integers(0, 2 ** 40, int(odd.sum()), dtype=np.uint64)
        return v
    if t == O.BOOL8:
        return rng.choice(np.array([0, 1, 2, 255], np.uint8), n)
    info = np.iinfo(O.NP[t])
    return rng.integers(info.min, info.max, n, dtype=O.NP[t], endpoint=True)


def _device(cols):
    return ca.Table([ca.Column.from_numpy(v, valid, t) for t, v, valid in cols])


TYPES = [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT8, O.UINT16, O.UINT32, O.UINT64, O.FLOAT32, O.FLOAT64, O.BOOL8,
         TS_DAYS, TS_MS]


@pytest.mark.parametrize("t", TYPES)
@pytest.mark.parametrize("seed", [0, 15, 42])
def test_single_column_matches_oracle(gpu, t, seed):
    rng = np.random.default_rng(1000 * t + seed)
    n = 20_000
    v = _values(rng, t, n)
    for valid in (None, rng.random(n) > 0.1):
        got, _ = hashing.murmurhash3_x86_32(_device([(t, v, valid)]), seed).to_numpy()
        want = O.murmurhash3_table([O.HostColumn(t, v, valid)], seed)
        assert np.array_equal(got.astype(np.uint32), want), f"type {t} seed {seed} nulls {valid is not None}"
        sgot, _ = hashing.spark_murmurhash3_x86_32(_device([(t, v, valid)]), seed).to_numpy()
        swant = O.spark_murmurhash3_table([O.HostColumn(t, v, valid)], seed)
        assert np.array_equal(sgot.astype(np.int32), swant), f"spark: type {t} seed {seed}"


@pytest.mark.parametrize("seed", [0, 15])
def test_composite_rows_match_oracle(gpu, seed):
    """Rows of up to eight columns of mixed types and nulls: hash_combine over
    the columns in order (row_operators.cuh:1868-1883)."""
    rng = np.random.default_rng(77 + seed)
    n = 50_000
    cols = []
    for t in (O.INT32, O.INT64, O.FLOAT64, O.BOOL8, O.UINT16, O.FLOAT32, O.INT8, TS_MS):
        valid = rng.random(n) > 0.15 if t in (O.INT64, O.FLOAT32, O.INT8) else None
        cols.append((t, _values(rng, t, n), valid))
    for k in (2, 5, 8):
        got, _ = hashing.murmurhash3_x86_32(_device(cols[:k]), seed).to_numpy()
        want = O.murmurhash3_table([O.HostColumn(t, v, valid) for t, v, valid in cols[:k]], seed)
        assert np.array_equal(got.astype(np.uint32), want), f"{k} columns"
        sgot, _ = hashing.spark_murmurhash3_x86_32(_device(cols[:k]), seed).to_numpy()
        swant = O.spark_murmurhash3_table([O.HostColumn(t, v, valid) for t, v, valid in cols[:k]], seed)
        assert np.array_equal(sgot.astype(np.int32), swant), f"spark: {k} columns"


def test_spark_golden_vectors(gpu):
    """cpp/tests/hashing/spark_murmurhash3_x86_32_test.cpp:221-318 (Apache
    Spark's values), one column at a time with seed 42."""
    spark = VECTORS["spark"]
    for name, case in spark.items():
        if name == "cite":
            continue
        t = {"INT32": O.INT32, "INT64": O.INT64, "INT16": O.INT16, "INT8": O.INT8, "BOOL8": O.BOOL8,
             "TIMESTAMP_DAYS": TS_DAYS, "TIMESTAMP_MILLISECONDS": TS_MS, "FLOAT32": O.FLOAT32,
             "FLOAT64": O.FLOAT64}[case["type"]]
        if "bits" in case:
            ut = np.uint32 if t == O.FLOAT32 else np.uint64
            v = np.array([int(b, 16) for b in case["bits"]], ut).view(O.NP[t])
        else:
            v = np.array(case["values"], dtype=np.int64).astype(O.NP[t])
        got, _ = hashing.spark_murmurhash3_x86_32(_device([(t, v, None)]), VECTORS["spark_seed"]).to_numpy()
        assert list(got.astype(np.int32)) == case["expect"], name


def test_reference_invariants(gpu):
    """murmurhash3_x86_32_test.cpp:31-121 without its string column: bool
    truth values 1, 2 and 255 hash alike, and null rows hash alike whatever
    their storage holds (0xFFFFFFFF per null element, :1811 of
    row_operators.cuh)."""
    imin, imax = np.iinfo(np.int32).min, np.iinfo(np.int32).max
    ints = np.array([0, 100, -100, imin, imax], np.int32)
    secs = np.array([0, 100, -100, np.iinfo(np.int64).min, np.iinfo(np.int64).max], np.int64)
    t1 = [(O.INT32, ints, None), (O.BOOL8, np.array([0, 1, 1, 1, 0], np.uint8), None), (13, secs, None)]
    t2 = [(O.INT32, ints, None), (O.BOOL8, np.array([0, 1, 2, 255, 0], np.uint8), None), (13, secs, None)]
    h1, _ = hashing.murmurhash3_x86_32(_device(t1)).to_numpy()
    h2, _ = hashing.murmurhash3_x86_32(_device(t2)).to_numpy()
    assert np.array_equal(h1, h2)
    iv = np.array([1, 0, 0, 1, 1], bool)
    bv = np.array([1, 1, 0, 0, 1], bool)
    n1 = [(O.INT32, ints, iv), (O.BOOL8, np.array([0, 1, 0, 1, 1], np.uint8), bv), (13, secs, iv)]
    n2 = [(O.INT32, np.array([0, -200, 200, imin, imax], np.int32), iv),
          (O.BOOL8, np.array([0, 2, 1, 0, 255], np.uint8), bv),
          (13, np.array([0, -200, 200, secs[3], secs[4]], np.int64), iv)]
    g1, _ = hashing.murmurhash3_x86_32(_device(n1)).to_numpy()
    g2, _ = hashing.murmurhash3_x86_32(_device(n2)).to_numpy()
    assert np.array_equal(g1, g2)
    # and the same rows through the oracle
    want = O.murmurhash3_table([O.HostColumn(t, v, valid) for t, v, valid in n1])
    assert np.array_equal(g1.astype(np.uint32), want)
    # -0.0 and +0.0, and every NaN, hash alike (hash_functions.cuh:29-47)
    f = np.array([0.0, -0.0, np.nan, -np.nan], np.float64)
    hf, _ = hashing.murmurhash3_x86_32(_device([(O.FLOAT64, f, None)])).to_numpy()
    assert hf[0] == hf[1] and hf[2] == hf[3] and hf[0] != hf[2]


def test_sliced_column_offset(gpu):
    """A column view with an offset hashes the rows from its offset (values
    and validity bits)."""
    rng = np.random.default_rng(5)
    n, off = 1000, 37
    v = rng.integers(-1000, 1000, n).astype(np.int64)
    valid = rng.random(n) > 0.3
    col = ca.Column.from_numpy(v, valid, O.INT64)
    sliced = ca.Column(col.type_id, n - off, col.data, col.mask, int((~valid[off:]).sum()), offset=off)
    got, _ = hashing.murmurhash3_x86_32(ca.Table([sliced])).to_numpy()
    want = O.murmurhash3_table([O.HostColumn(O.INT64, v[off:], valid[off:])])
    assert np.array_equal(got.astype(np.uint32), want)
