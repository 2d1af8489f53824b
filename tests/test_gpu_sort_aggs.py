"""GPU parity of the aggregations cudf serves from its sort-based groupby
(MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT) through the C ABI, against the
oracle on seeded inputs: every value type, each interpolation, both null
policies, NaN and ties, negative and out-of-range n, composite and null
keys, each requested execution (the partitioned one gives way to the global
one), and sizes up to 2M rows. Results must match exactly, except float
quantiles between two values (within 1e-12 relative, the interpolation's
rounding)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import gpu_groupby, oracle_groupby, sorted_result

pytestmark = pytest.mark.gpu


def Q(q, interp=0):
    return (O.QUANTILE, 0, 0, interp, q)


def NTH(n, include=True):
    return (O.NTH_ELEMENT, n, 1 if include else 0, 0, 0.5)


def NU(include=False):
    return (O.NUNIQUE, 0, 1 if include else 0, 0, 0.5)


MED = (O.MEDIAN, 0, 0, 0, 0.5)
MIX = [MED, Q(0.25), Q(0.9, 1), Q(0.33, 2), Q(0.5, 3), Q(0.7, 4), NU(), NU(True), NTH(0), NTH(-1), NTH(2, False),
       NTH(-3, False), NTH(50)]


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
    for r, (gblk, oblk) in enumerate(zip(gr, orr)):
        for a, ((g, gvalid), (o, ovalid)) in zip(reqs[r][1], zip(gblk, oblk)):
            what = f"request {r} agg {a}"
            gm = np.ones(len(g), bool) if gvalid is None else np.asarray(gvalid, bool)
            om = np.ones(len(o), bool) if ovalid is None else np.asarray(ovalid, bool)
            assert np.array_equal(gm, om), f"{what}: validity differs at {np.flatnonzero(gm != om)[:5]}"
            g, o = np.asarray(g)[om], np.asarray(o)[om]
            if g.dtype.kind == "f":
                with np.errstate(invalid="ignore"):  # inf - inf where both are inf
                    same = (g == o) | (np.isnan(g) & np.isnan(o)) | (np.abs(g - o) <= 1e-12 * np.abs(o))
                assert same.all(), f"{what}: {int((~same).sum())} groups differ, e.g. {g[~same][:3]} vs {o[~same][:3]}"
            else:
                assert np.array_equal(g, o), f"{what}: differs at {np.flatnonzero(g != o)[:5]}"


def _values(rng, vt, n, ties=False):
    if vt in (O.FLOAT32, O.FLOAT64):
        if ties:
            pool = np.array([-np.inf, -1.5, -0.0, 0.0, 2.5, np.inf, np.nan], dtype=O.NP[vt])
            return pool[rng.integers(0, len(pool), n)]
        return (rng.standard_normal(n) * 50).astype(O.NP[vt])
    info = np.iinfo(O.NP[vt])
    if ties:
        return rng.integers(max(info.min, -4), min(info.max, 4), n, endpoint=True).astype(O.NP[vt])
    return rng.integers(max(info.min, -10 ** 9), min(info.max, 10 ** 9), n, endpoint=True).astype(O.NP[vt])


@pytest.mark.parametrize("vt", [O.INT8, O.INT16, O.INT32, O.INT64, O.UINT8, O.UINT32, O.UINT64, O.FLOAT32,
                                O.FLOAT64])
def test_value_types(gpu, vt):
    rng = np.random.default_rng(700 + vt)
    n = 50_000
    keys = rng.integers(0, 3000, n).astype(np.int32)
    reqs = [((vt, _values(rng, vt, n), rng.random(n) > 0.1), MIX)]
    _compare([(O.INT32, keys, None)], reqs)


@pytest.mark.parametrize("vt", [O.INT16, O.FLOAT32, O.FLOAT64])
def test_ties_and_nan(gpu, vt):
    rng = np.random.default_rng(720 + vt)
    n = 60_000
    keys = rng.integers(0, 500, n).astype(np.int64)
    _compare([(O.INT64, keys, None)], [((vt, _values(rng, vt, n, ties=True), rng.random(n) > 0.05), MIX)])


PATHS = ["auto", "partitioned", "direct", "global"]


@pytest.mark.parametrize("path", PATHS)
@pytest.mark.parametrize("card", [7, 20_000])
def test_every_execution(gpu, monkeypatch, path, card):
    if path != "auto":
        monkeypatch.setenv("MGB_GROUPBY_PATH", path)
    rng = np.random.default_rng(740 + card)
    n = 300_000
    keys = rng.integers(-2 ** 40, 2 ** 40, card)[rng.integers(0, card, n)]
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.02), [MED, Q(0.1), NU(True), NTH(-2)]),
            ((O.INT32, _values(rng, O.INT32, n, ties=True), None), [NU(), NTH(1, False), Q(0.5, 3)])]
    _compare([(O.INT64, keys, None)], reqs)


def test_composite_and_null_keys(gpu):
    rng = np.random.default_rng(760)
    n = 80_000
    a = rng.integers(0, 100, n).astype(np.int32)
    b = rng.integers(0, 20, n).astype(np.float64) / 2
    keys = [(O.INT32, a, rng.random(n) > 0.05), (O.FLOAT64, b, None)]
    reqs = [((O.INT64, _values(rng, O.INT64, n), rng.random(n) > 0.1), [MED, Q(0.3, 4), NU(True), NTH(0, False)])]
    _compare(keys, reqs)
    _compare(keys, reqs, include_null_keys=True)


def test_mixed_with_hash_aggregations(gpu):
    rng = np.random.default_rng(770)
    n = 100_000
    keys = rng.integers(0, 8000, n).astype(np.int32)
    v = rng.standard_normal(n)
    valid = rng.random(n) > 0.1
    reqs = [((O.FLOAT64, v, valid), [O.SUM, MED, (O.VARIANCE, 1), NU(), O.ARGMAX, NTH(-1), O.COUNT_VALID])]
    gk, gr = gpu_groupby([(O.INT32, keys, None)], reqs)
    ok_, orr = oracle_groupby([(O.INT32, keys, None)], reqs)
    gk, gr = sorted_result(gk, gr)
    ok_, orr = sorted_result(ok_, orr)
    for (g, gv), (o, ov), a in zip(gr[0], orr[0], reqs[0][1]):
        gm = np.ones(len(g), bool) if gv is None else np.asarray(gv, bool)
        om = np.ones(len(o), bool) if ov is None else np.asarray(ov, bool)
        assert np.array_equal(gm, om), a
        assert np.allclose(np.asarray(g, np.float64)[om], np.asarray(o, np.float64)[om], rtol=1e-9, equal_nan=True), a


SEGMENT_SIZES = [1, 2, 63, 64, 65, 127, 128, 129, 255, 256, 257, 1000, 4095, 4096, 4097, 8192, 8193, 12_289, 40_000,
                 1_500_000]


@pytest.mark.parametrize("vt", [O.INT32, O.FLOAT32, O.FLOAT64, O.INT64])
def test_segment_sizes(gpu, vt):
    """Groups at every size boundary of the segmented sort: one wave (up to
    64, 128 and 256 keys), one workgroup (up to 4096) and chunked merges (one
    to nine rounds), rows shuffled; value orderings (narrow and wide codes)
    and both row orderings."""
    rng = np.random.default_rng(800 + vt)
    sizes = np.array(SEGMENT_SIZES)
    keys = np.repeat(np.arange(len(sizes), dtype=np.int64) * 7919 - 50_000, sizes)
    rng.shuffle(keys)
    n = keys.size
    for ties in (False, True):
        v = _values(rng, vt, n, ties=ties)
        reqs = [((vt, v, rng.random(n) > 0.03), [MED, Q(0.75, 1), Q(0.2), NU(), NU(True), NTH(0), NTH(-1), NTH(-5, False),
                                                 NTH(4000, False)])]
        _compare([(O.INT64, keys, None)], reqs)


@pytest.mark.parametrize("card", [1, 2, 127, 128, 255, 256, 16_383, 16_384, 65_536])
def test_radix_pass_boundaries(gpu, card):
    """Group counts at the edges of the radix passes over the group bits
    (one pass of up to 8 bits, two, three with uneven digit widths), every
    group present, value and row orderings."""
    rng = np.random.default_rng(860 + card)
    n = max(4 * card, 50_000)
    keys = np.concatenate([np.arange(card), rng.integers(0, card, n - card)]).astype(np.int64) * 3 - 7
    rng.shuffle(keys)
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.05), [MED, Q(0.9, 1), NU(True), NTH(0), NTH(-1, False)]),
            ((O.INT32, _values(rng, O.INT32, n, ties=True), rng.random(n) > 0.05), [MED, NU(), NTH(1)])]
    _compare([(O.INT64, keys, None)], reqs)


def test_large(gpu):
    rng = np.random.default_rng(780)
    n = 2_000_000
    keys = rng.integers(0, 200_000, n).astype(np.int64) * 31
    reqs = [((O.FLOAT64, _values(rng, O.FLOAT64, n), rng.random(n) > 0.05), [MED, Q(0.95), NU(), NTH(3)])]
    _compare([(O.INT64, keys, None)], reqs)


def test_twenty_million_rows_against_numpy(gpu):
    """At 20M rows over 1M int64 keys (about 20 rows a group), each group's
    sorted valid values from numpy: MEDIAN (linear), QUANTILE
    0.9 LOWER, NUNIQUE, the first row (INCLUDE) and the last valid value
    (EXCLUDE)."""
    rng = np.random.default_rng(790)
    n, card = 20_000_000, 1_000_000
    keys = rng.integers(-2 ** 62, 2 ** 62, card)[rng.integers(0, card, n)]
    v = np.round(rng.standard_normal(n) * 100, 1)  # repeats, so NUNIQUE < count
    valid = rng.random(n) > 0.05
    reqs = [((O.FLOAT64, v, valid), [MED, Q(0.9, 1), NU(), NTH(0), NTH(-1, False)])]
    gk, gr = gpu_groupby([(O.INT64, keys, None)], reqs)
    order = np.argsort(gk[0][0], kind="stable")
    got = [(np.asarray(a)[order], np.ones(len(order), bool) if m is None else np.asarray(m, bool)[order])
           for a, m in gr[0]]

    ukeys, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    assert np.array_equal(np.asarray(gk[0][0])[order], ukeys)
    G = len(ukeys)
    vi = np.flatnonzero(valid)
    srt = np.lexsort((v[vi], inv[vi]))
    sg, sv = inv[vi][srt], v[vi][srt]
    starts = np.searchsorted(sg, np.arange(G))
    cnt = np.searchsorted(sg, np.arange(G), "right") - starts
    has = cnt > 0
    pos = 0.5 * (cnt[has] - 1)
    lo, hi = np.floor(pos).astype(np.int64), np.ceil(pos).astype(np.int64)
    f = pos - lo
    med = (1.0 - f) * sv[starts[has] + lo] + f * sv[starts[has] + hi]
    (gm, gmv), (gq, gqv), (gu, guv), (g0, g0v), (gl, glv) = got
    assert np.array_equal(gmv, has)
    assert np.all(np.abs(gm[has] - med) <= 1e-12 * np.abs(med)), "MEDIAN"
    q_lo = np.floor(0.9 * (cnt[has] - 1)).astype(np.int64)
    assert np.array_equal(gqv, has) and np.array_equal(gq[has], sv[starts[has] + q_lo]), "QUANTILE LOWER"
    d = np.ones(len(sv), bool)
    d[1:] = (sg[1:] != sg[:-1]) | (sv[1:] != sv[:-1])
    assert np.array_equal(gu, np.bincount(sg, weights=d, minlength=G).astype(np.int64)), "NUNIQUE"
    assert np.array_equal(g0v, valid[first]) and np.array_equal(g0[g0v], v[first][g0v]), "NTH_ELEMENT 0"
    last = np.full(G, -1)
    rev_g, rev_i = np.unique(inv[vi][::-1], return_index=True)
    last[rev_g] = vi[len(vi) - 1 - rev_i]
    assert np.array_equal(glv, last >= 0) and np.array_equal(gl[glv], v[last[last >= 0]]), "NTH_ELEMENT -1"
