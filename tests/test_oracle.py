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


def test_float_min_is_an_outcome_of_the_reference_fold():
    """cudf folds min(old, x) = old < x ? old : x from +inf in the order its
    atomics land; the oracle's MIN is that fold over the NaN rows first."""
    def fold(xs):
        old = np.inf
        for x in xs:
            old = old if old < x else x
        return old

    keys = np.zeros(3, np.int32)
    for v in ([1.0, np.nan, 0.5], [1.0, 2.0, np.nan], [np.nan, np.nan, np.nan], [np.nan, 3.0, -2.0]):
        v = np.array(v)
        _, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(O.FLOAT64, v), [O.MIN])])
        want = fold(np.concatenate([v[np.isnan(v)], v[~np.isnan(v)]]))
        got = res[0][0][0][0]
        assert (math.isnan(got) and math.isnan(want)) or got == want, (v, got, want)


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


# ---------------------------------------------------------------- second-pass aggregations


def _tiny_and_large_groups(rng, n):
    """int32 keys: about 300 groups of ~17 rows plus groups of 1 and 2 rows."""
    keys = rng.integers(0, 300, n).astype(np.int32)
    keys[:20] = np.repeat(np.arange(1000, 1010, dtype=np.int32), 2)
    keys[20:30] = np.arange(2000, 2010, dtype=np.int32)
    return keys


@pytest.mark.parametrize("vt", [O.INT8, O.INT32, O.INT64, O.UINT16, O.FLOAT32, O.FLOAT64])
@pytest.mark.parametrize("ddof", [0, 1, 2, 7])
def test_oracle_variance_std_against_numpy(vt, ddof):
    """VARIANCE / STD per group = numpy.var / numpy.std with the same ddof;
    valid exactly when the group has more valid values than ddof."""
    rng = np.random.default_rng(40 + vt * 10 + ddof)
    n = 5000
    keys = _tiny_and_large_groups(rng, n)
    if vt in (O.FLOAT32, O.FLOAT64):
        vals = (rng.standard_normal(n) * 100 + 7).astype(O.NP[vt])
    else:
        info = np.iinfo(O.NP[vt])
        vals = rng.integers(max(info.min, -10 ** 6), min(info.max, 10 ** 6), n, endpoint=True).astype(O.NP[vt])
    valid = rng.random(n) > 0.2
    ks, res = O.groupby([O.HostColumn(O.INT32, keys)],
                        [(O.HostColumn(vt, vals, valid), [(O.VARIANCE, ddof), (O.STD, ddof)])])
    (var, var_ok), (std, std_ok) = res[0]
    for g, k in enumerate(ks[0][0]):
        x = vals[(keys == k) & valid].astype(np.float64)
        ok = len(x) > 0 and len(x) > ddof
        assert bool(var_ok[g]) == ok and bool(std_ok[g]) == ok, (k, len(x), ddof)
        if not ok:
            continue
        # float32 sources take the mean from a float32 sum, as cudf does
        rtol = 1e-9 if vt == O.FLOAT32 else 1e-12
        scale = float(np.mean(x * x)) * len(x) / (len(x) - ddof)
        assert abs(var[g] - np.var(x, ddof=ddof)) <= rtol * scale, (k, var[g], np.var(x, ddof=ddof))
        assert abs(std[g] - np.std(x, ddof=ddof)) <= rtol * np.sqrt(scale) + 1e-300, (k, std[g])


def _brute_arg(x, rows, is_min, is_float):
    """Set-based ARGMIN / ARGMAX: the smallest row index among the rows holding
    the extreme value; NaN only when every value is NaN; -0.0 below +0.0."""
    if is_float:
        nan = np.isnan(x)
        if nan.all():
            return rows[0]
        x, rows = x[~nan], rows[~nan]
    best = x.min() if is_min else x.max()
    cand = x == best
    if is_float and best == 0:
        neg = np.signbit(x) & cand
        pos = ~np.signbit(x) & cand
        if is_min and neg.any():
            cand = neg
        elif not is_min and pos.any():
            cand = pos
    return rows[cand].min()


@pytest.mark.parametrize("vt", [O.INT8, O.INT32, O.UINT32, O.INT64, O.UINT64, O.FLOAT32, O.FLOAT64, O.BOOL8])
def test_oracle_argmin_argmax_brute_force(vt):
    rng = np.random.default_rng(60 + vt)
    n = 4000
    keys = _tiny_and_large_groups(rng, n)
    if vt in (O.FLOAT32, O.FLOAT64):
        pool = np.array([-np.inf, -1.5, -0.0, 0.0, 2.5, np.inf, np.nan], dtype=O.NP[vt])
        vals = pool[rng.integers(0, len(pool), n)]
        vals[keys == 1003] = np.nan  # an all-NaN group
    elif vt == O.BOOL8:
        vals = rng.integers(0, 2, n).astype(np.uint8)
    else:
        info = np.iinfo(O.NP[vt])
        pool = np.array([info.min, info.min + 1, 0, 1, 2, info.max], dtype=O.NP[vt])  # many ties
        vals = pool[rng.integers(0, len(pool), n)]
    valid = rng.random(n) > 0.15
    ks, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(vt, vals, valid), [O.ARGMIN, O.ARGMAX])])
    (amin, amin_ok), (amax, amax_ok) = res[0]
    rows = np.arange(n)
    is_float = vt in (O.FLOAT32, O.FLOAT64)
    for g, k in enumerate(ks[0][0]):
        sel = (keys == k) & valid
        assert bool(amin_ok[g]) == sel.any() and bool(amax_ok[g]) == sel.any()
        if not sel.any():
            continue
        assert amin[g] == _brute_arg(vals[sel], rows[sel], True, is_float), (k, amin[g])
        assert amax[g] == _brute_arg(vals[sel], rows[sel], False, is_float), (k, amax[g])


@pytest.mark.parametrize("vt", [O.FLOAT32, O.FLOAT64])
def test_oracle_float_min_max_nan_and_zero(vt):
    """MIN / MAX over floats: numpy's nanmin / nanmax per group, NaN for a
    group of NaN only, and the signed zero -0.0 for MIN / +0.0 for MAX where
    both zeros are the extreme."""
    rng = np.random.default_rng(80 + vt)
    n = 4000
    keys = _tiny_and_large_groups(rng, n)
    pool = np.array([-np.inf, -1.5, -0.0, 0.0, 2.5, np.inf, np.nan], dtype=O.NP[vt])
    vals = pool[rng.integers(0, len(pool), n)]
    vals[keys == 1004] = np.nan
    valid = rng.random(n) > 0.1
    ks, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(vt, vals, valid), [O.MIN, O.MAX])])
    (mn, mn_ok), (mx, mx_ok) = res[0]
    for g, k in enumerate(ks[0][0]):
        x = vals[(keys == k) & valid]
        assert bool(mn_ok[g]) == (len(x) > 0) and bool(mx_ok[g]) == (len(x) > 0)
        if not len(x):
            continue
        if np.isnan(x).all():
            assert np.isnan(mn[g]) and np.isnan(mx[g])
            continue
        lo, hi = np.nanmin(x), np.nanmax(x)
        assert mn[g] == lo and mx[g] == hi, (k, mn[g], mx[g], lo, hi)
        if lo == 0:
            assert np.signbit(mn[g]) == np.signbit(x[x == 0]).any()
        if hi == 0:
            assert np.signbit(mx[g]) == np.signbit(x[x == 0]).all()


# ---------------------------------------------------------------- sort-path aggregations

NP_METHOD = {0: "linear", 1: "lower", 2: "higher", 3: "midpoint", 4: "nearest"}


@pytest.mark.parametrize("vt", [O.INT16, O.INT64, O.UINT32, O.FLOAT32, O.FLOAT64])
@pytest.mark.parametrize("interp", [0, 1, 2, 3, 4])
def test_oracle_quantiles_against_numpy(vt, interp):
    """QUANTILE / MEDIAN per group = numpy.quantile of the group's valid
    values with the matching method; null for a group without values."""
    rng = np.random.default_rng(90 + vt * 7 + interp)
    n = 3000
    keys = _tiny_and_large_groups(rng, n)
    if vt in (O.FLOAT32, O.FLOAT64):
        vals = (rng.standard_normal(n) * 100).astype(O.NP[vt])
    else:
        info = np.iinfo(O.NP[vt])
        vals = rng.integers(max(info.min, -10 ** 6), min(info.max, 10 ** 6), n, endpoint=True).astype(O.NP[vt])
    valid = rng.random(n) > 0.2
    qs = [0.0, 0.13, 0.5, 0.77, 1.0]
    aggs = [(O.QUANTILE, 0, 0, interp, q) for q in qs] + [(O.MEDIAN, 0, 0, 0, 0.5)]
    ks, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(vt, vals, valid), aggs)])
    for g, k in enumerate(ks[0][0]):
        x = np.sort(vals[(keys == k) & valid].astype(np.float64))
        for a, (out, ok) in zip(aggs, res[0]):
            assert bool(ok[g]) == (len(x) > 0)
            if not len(x):
                continue
            method = NP_METHOD[a[3]]
            want = np.quantile(x, a[4], method=method)
            assert abs(out[g] - want) <= 1e-9 * max(1.0, abs(want)), (k, a, out[g], want)


@pytest.mark.parametrize("vt", [O.INT8, O.INT32, O.FLOAT64])
def test_oracle_nunique_and_nth_against_python(vt):
    rng = np.random.default_rng(95 + vt)
    n = 3000
    keys = _tiny_and_large_groups(rng, n)
    if vt == O.FLOAT64:
        vals = rng.choice(np.array([-0.0, 0.0, 1.5, np.nan, 2.5]), n)
    else:
        vals = rng.integers(-3, 4, n).astype(O.NP[vt])
    valid = rng.random(n) > 0.25
    aggs = [(O.NUNIQUE, 0, 0, 0, 0.5), (O.NUNIQUE, 0, 1, 0, 0.5)] + \
           [(O.NTH_ELEMENT, m, p, 0, 0.5) for m in (0, 2, -1, -3) for p in (0, 1)]
    ks, res = O.groupby([O.HostColumn(O.INT32, keys)], [(O.HostColumn(vt, vals, valid), aggs)])
    for g, k in enumerate(ks[0][0]):
        rows = np.flatnonzero(keys == k)
        vv = [float(vals[r]) for r in rows if valid[r]]
        distinct = set()
        for x in vv:  # NaN == NaN, -0.0 == +0.0
            distinct.add("nan" if np.isnan(x) else x + 0.0)
        for a, (out, ok) in zip(aggs, res[0]):
            if a[0] == O.NUNIQUE:
                want = len(distinct) + (1 if a[2] == 1 and (~valid[rows]).any() else 0)
                assert out[g] == want and ok[g], (k, a, out[g], want)
            else:
                pool = list(rows) if a[2] == 1 else [r for r in rows if valid[r]]
                idx = a[1] if a[1] >= 0 else len(pool) + a[1]
                hit = 0 <= idx < len(pool) and valid[pool[idx]]
                assert bool(ok[g]) == hit, (k, a)
                if hit:
                    assert out[g] == vals[pool[idx]] or (np.isnan(out[g]) and np.isnan(vals[pool[idx]])), (k, a)
