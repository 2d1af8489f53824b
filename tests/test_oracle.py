"""The oracle pinned to the reference: SMHasher's MurmurHash3_x86_32
verification value, the Spark and cudf golden hash vectors, the literal
fixtures of the reference's groupby tests, brute force on small inputs and
pandas (the cudf.pandas CPU fallback) on random inputs."""
from __future__ import annotations

import json
import math
import os
import struct

import numpy as np
import pytest

from oracle import oracle as O
from tests import golden_cases as G

HERE = os.path.dirname(os.path.abspath(__file__))
HASH = json.load(open(os.path.join(HERE, "golden", "hash_vectors.json")))
TYPE = {"INT8": O.INT8, "INT16": O.INT16, "INT32": O.INT32, "INT64": O.INT64, "BOOL8": O.BOOL8,
        "FLOAT32": O.FLOAT32, "FLOAT64": O.FLOAT64, "TIMESTAMP_DAYS": 12, "TIMESTAMP_MILLISECONDS": 14}


def test_smhasher_verification_value():
    key = bytes(range(256))
    hashes = b"".join(struct.pack("<I", O.murmur3(key[:i], 256 - i)) for i in range(256))
    assert O.murmur3(hashes, 0) == HASH["smhasher_verification"]["value"]


def test_hash_combine_closed_form():
    for lhs, rhs in [(0, 0), (1, 2), (0xFFFFFFFF, 0x12345678), (0xDEADBEEF, 0)]:
        expect = (lhs ^ ((rhs + 0x9E3779B9 + ((lhs << 6) & 0xFFFFFFFF) + (lhs >> 2)) & 0xFFFFFFFF)) & 0xFFFFFFFF
        assert O.hash_combine(lhs, rhs) == expect


def spark_column(spec):
    t = TYPE[spec["type"]]
    if "bits" in spec:
        raw = np.array([int(b, 16) for b in spec["bits"]], dtype=np.uint32 if t == O.FLOAT32 else np.uint64)
        return O.HostColumn(t, raw.view(O.NP[t]))
    return O.HostColumn(t, np.array(spec["values"], dtype=np.int64).astype(O.NP[t]))


@pytest.mark.parametrize("name", sorted(k for k in HASH["spark"] if k != "cite"))
def test_spark_murmur_golden(name):
    spec = HASH["spark"][name]
    got = O.spark_murmurhash3_table([spark_column(spec)], HASH["spark_seed"])
    assert got.tolist() == spec["expect"]


@pytest.mark.parametrize("seed", [0, 15])
def test_cudf_list_hash_golden(seed):
    # pins murmur3 over 4- and 8-byte elements and hash_combine together
    d = HASH["cudf_list"]
    out = []
    for row in d["rows"]:
        inner = O.hash_combine(0, O.murmur3(struct.pack("<i", len(row)), 0))
        for e in row:
            inner = O.hash_combine(inner, O.murmur3(struct.pack("<Q", e), seed))
        out.append(O.hash_combine(seed, inner))
    assert out == [v & 0xFFFFFFFF for v in d[f"seed{seed}"]]


def test_row_hash_is_combine_of_column_hashes():
    a = np.array([1, -7, 123456], dtype=np.int32)
    b = np.array([2.5, -0.0, np.nan], dtype=np.float64)
    got = O.murmurhash3_table([O.HostColumn(O.INT32, a), O.HostColumn(O.FLOAT64, b)], 0)
    for i in range(3):
        hb = 0.0 if b[i] == 0 else (float("nan") if math.isnan(b[i]) else float(b[i]))
        eb = struct.pack("<d", hb) if not math.isnan(hb) else struct.pack("<Q", 0x7FF8000000000000)
        h = O.hash_combine(0, O.murmur3(struct.pack("<i", int(a[i])), 0))
        h = O.hash_combine(h, O.murmur3(eb, 0))
        assert got[i] == h


def test_hash_invariants_nulls_and_zero_and_nan():
    # murmurhash3_x86_32_test.cpp:31-121: nulls hash alike whatever their payload;
    # -0.0 hashes like +0.0; any NaN like the canonical NaN
    vals = np.array([5, 6], dtype=np.int64)
    h = O.murmurhash3_table([O.HostColumn(O.INT64, vals, np.array([False, False]))], 0)
    assert h[0] == h[1] == O.hash_combine(0, 0xFFFFFFFF)
    f = np.array([0.0, -0.0], dtype=np.float32)
    hf = O.murmurhash3_table([O.HostColumn(O.FLOAT32, f)], 0)
    assert hf[0] == hf[1]
    n = np.array([0x7FC00000, 0xFFC00000, 0x7F800001], dtype=np.uint32).view(np.float32)
    hn = O.murmurhash3_table([O.HostColumn(O.FLOAT32, n)], 0)
    assert hn[0] == hn[1] == hn[2]


CASES = G.load_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["id"] for c in CASES])
def test_reference_fixture_through_oracle(case):
    kt, kv, kvalid = case["keys"]
    keys, res = O.groupby([O.HostColumn(kt, kv, kvalid)], [(O.HostColumn(*case["vals"]), [case["kind"]])],
                          case["include_null_keys"])
    G.check_case(case, keys, res)


def brute_force(keys, vals, valid, kind):
    groups = {}
    for k, v, ok in zip(keys.tolist(), vals.tolist(), valid.tolist()):
        groups.setdefault(k, []).append(v if ok else None)
    out = {}
    for k, vs in groups.items():
        xs = [x for x in vs if x is not None]
        if kind == O.COUNT_VALID:
            out[k] = (len(xs), True)
        elif kind == O.COUNT_ALL:
            out[k] = (len(vs), True)
        elif not xs:
            out[k] = (None, False)
        elif kind == O.SUM:
            out[k] = (sum(xs), True)
        elif kind == O.PRODUCT:
            out[k] = (math.prod(xs), True)
        elif kind == O.SUM_OF_SQUARES:
            out[k] = (sum(x * x for x in xs), True)
        elif kind == O.MIN:
            out[k] = (min(xs), True)
        elif kind == O.MAX:
            out[k] = (max(xs), True)
        elif kind == O.MEAN:
            out[k] = (sum(xs) / len(xs), True)
    return out


@pytest.mark.parametrize("kind", [O.SUM, O.PRODUCT, O.SUM_OF_SQUARES, O.MIN, O.MAX, O.COUNT_VALID, O.COUNT_ALL,
                                  O.MEAN])
def test_oracle_vs_brute_force(kind):
    rng = np.random.default_rng(kind)
    n = 300
    keys = rng.integers(0, 12, n).astype(np.int32)
    vals = rng.integers(-5, 6, n).astype(np.int64)
    valid = rng.random(n) > 0.2
    ks, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(O.INT64, vals, valid), [kind])])
    bf = brute_force(keys, vals, valid, kind)
    got_v, got_ok = res[0][0]
    assert len(ks[0][0]) == len(bf)
    for k, v, ok in zip(ks[0][0].tolist(), got_v.tolist(), got_ok.tolist()):
        want, want_ok = bf[k]
        assert ok == want_ok
        if ok:
            if kind == O.MEAN:
                assert abs(v - want) <= 1e-12 * max(1.0, abs(want))
            else:
                # int64 results wrap like the device's two's complement arithmetic
                assert v == ((want + 2 ** 63) % 2 ** 64) - 2 ** 63


def test_int32_sum_widens_to_int64():
    ks, res = O.groupby([O.HostColumn(O.INT32, np.zeros(3, np.int32))],
                        [(O.HostColumn(O.INT32, np.full(3, 2 ** 31 - 1, np.int32)), [O.SUM])])
    assert res[0][0][0][0] == 3 * (2 ** 31 - 1)


def test_float_min_follows_sequential_cas_semantics():
    # min(old, x) = old < x ? old : x applied in row order from +inf
    keys = np.zeros(3, np.int32)
    v = np.array([1.0, np.nan, 0.5])
    _, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(O.FLOAT64, v), [O.MIN])])
    assert res[0][0][0][0] == 0.5
    _, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(O.FLOAT64, np.array([1.0, 2.0, np.nan])), [O.MIN])])
    assert math.isnan(res[0][0][0][0])


def test_target_types_closed_form():
    # cpp/include/cudf/detail/aggregation/aggregation.hpp:1168-1424
    assert O.target_type(O.INT8, O.SUM) == O.INT64
    assert O.target_type(O.UINT32, O.SUM) == O.INT64
    assert O.target_type(O.FLOAT32, O.SUM) == O.FLOAT32
    assert O.target_type(O.INT16, O.MIN) == O.INT16
    assert O.target_type(O.FLOAT64, O.COUNT_VALID) == O.INT32
    assert O.target_type(O.INT32, O.MEAN) == O.FLOAT64
    assert O.target_type(12, O.SUM) == -1           # timestamps cannot be summed
    assert O.target_type(12, O.MIN) == 12
    assert O.target_type(18, O.MEAN) == 18          # duration mean stays a duration


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_vs_pandas(seed):
    rng = np.random.default_rng(seed)
    n = 50_000
    keys = rng.integers(-300, 300, n).astype(np.int64)
    cols = [(rng.integers(-1_000_000, 1_000_001, n).astype(np.int32), rng.random(n) > 0.05),
            (rng.standard_normal(n), rng.random(n) > 0.05),
            (rng.standard_normal(n).astype(np.float32), rng.random(n) > 0.05)]
    kinds = [O.SUM, O.COUNT_VALID, O.MIN, O.MAX, O.MEAN]
    tmap = {np.dtype(np.int32): O.INT32, np.dtype(np.float64): O.FLOAT64, np.dtype(np.float32): O.FLOAT32}
    ok, ores = O.groupby([O.HostColumn(O.INT64, keys)], [(O.HostColumn(tmap[v.dtype], v, m), kinds) for v, m in cols])
    order = O.sort_order(ok)
    pk, pres = O.pandas_groupby([(keys, None)], cols, [kinds] * 3)
    assert np.array_equal(ok[0][0][order], pk[0][0])
    for (v, m), oblk, pblk in zip(cols, ores, pres):
        absum = O.groupby([O.HostColumn(O.INT64, keys)],
                          [(O.HostColumn(O.FLOAT64, np.abs(v.astype(np.float64)), m), [O.SUM])])[1][0][0][0][order]
        for kind, (a, av), (b, bv) in zip(kinds, oblk, pblk):
            a, av = a[order], av[order]
            assert np.array_equal(av, bv)
            a, b = a[av], b[bv]
            if v.dtype.kind == "f" and kind in (O.SUM, O.MEAN):
                tol = 1e-5 if v.dtype == np.float32 else 1e-12
                scale = absum[av] / (np.maximum(1, (kind == O.MEAN) * oblk[1][0][order][av]))
                err = np.abs(a.astype(np.float64) - b.astype(np.float64))
                assert (err <= tol * np.maximum(np.abs(b), scale)).all()
            else:
                assert np.array_equal(a, b)
