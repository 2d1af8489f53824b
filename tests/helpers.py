"""Shared helpers of the parity tests: run a groupby through the C ABI and
through the oracle, sort both by key and compare at the stated bar
(integers bit-exact; float64 rel 1e-12, float32 rel 1e-5, each also allowed
the same error scaled by the group's sum of |x|)."""
from __future__ import annotations

import numpy as np

import cudf_amd as ca
from oracle import oracle as O

KIND_NAMES = {"SUM": O.SUM, "PRODUCT": O.PRODUCT, "MIN": O.MIN, "MAX": O.MAX, "COUNT_VALID": O.COUNT_VALID,
              "COUNT_ALL": O.COUNT_ALL, "SUM_OF_SQUARES": O.SUM_OF_SQUARES, "MEAN": O.MEAN}
TYPE_NAMES = {"INT8": O.INT8, "INT16": O.INT16, "INT32": O.INT32, "INT64": O.INT64, "UINT8": O.UINT8,
              "UINT16": O.UINT16, "UINT32": O.UINT32, "UINT64": O.UINT64, "FLOAT32": O.FLOAT32,
              "FLOAT64": O.FLOAT64, "BOOL8": O.BOOL8}


def to_agg(kind) -> ca.aggregation.Aggregation:
    """kind, or (kind, ddof) as the oracle takes it."""
    if isinstance(kind, tuple) and len(kind) >= 5:
        return ca.aggregation.Aggregation(ca.Kind(kind[0]), kind[1], kind[2], kind[3], kind[4])
    if isinstance(kind, tuple):
        return ca.aggregation.Aggregation(ca.Kind(kind[0]), kind[1])
    return ca.aggregation.Aggregation(ca.Kind(kind), 1 if kind in (O.VARIANCE, O.STD) else 0)


def gpu_groupby(keys, requests, include_null_keys=False, device="cuda"):
    """keys: [(type_id, values, valid)], requests: [((type_id, values, valid), [kind...])]
    -> (key columns [(values, valid)], results [[(values, valid)]]) from the HIP path."""
    kcols = [ca.Column.from_numpy(np.asarray(v), valid, t, device=device) for t, v, valid in keys]
    reqs = []
    for (t, v, valid), kinds in requests:
        col = ca.Column.from_numpy(np.asarray(v), valid, t, device=device)
        reqs.append(ca.GroupByRequest(col, [to_agg(k) for k in kinds]))
    gb = ca.GroupBy(ca.Table(kcols), ca.NullPolicy.INCLUDE if include_null_keys else ca.NullPolicy.EXCLUDE)
    out_keys, results = gb.aggregate(reqs)
    ok = [c.to_numpy() for c in out_keys.columns]
    res = [[c.to_numpy() for c in t.columns] for t in results]
    return ok, res


def oracle_groupby(keys, requests, include_null_keys=False):
    kcols = [O.HostColumn(t, v, valid) for t, v, valid in keys]
    reqs = [(O.HostColumn(t, v, valid), list(kinds)) for (t, v, valid), kinds in requests]
    return O.groupby(kcols, reqs, include_null_keys)


def sorted_result(key_cols, results):
    order = O.sort_order(key_cols)
    ks = [(v[order], None if valid is None else valid[order]) for v, valid in key_cols]
    rs = [[(v[order], None if valid is None else np.asarray(valid, bool)[order]) for v, valid in blk]
          for blk in results]
    return ks, rs


def abs_sums(keys, requests, include_null_keys=False):
    """Per request column: sum of |x| per group (oracle order), for the scaled float tolerance."""
    out = []
    for (t, v, valid), kinds in requests:
        vv = np.abs(np.asarray(v, dtype=np.float64))
        vv = np.where(np.isfinite(vv), vv, 0.0)
        kcols = [O.HostColumn(kt, kv, kvalid) for kt, kv, kvalid in keys]
        ks, res = O.groupby(kcols, [(O.HostColumn(O.FLOAT64, vv, valid), [O.SUM])], include_null_keys)
        out.append((ks, res[0][0][0]))
    return out


def assert_same(got, want, type_id_of_result, abs_sum=None, rtol=None, what=""):
    gv, gvalid = got
    wv, wvalid = want
    gvalid = np.ones(len(gv), bool) if gvalid is None else np.asarray(gvalid, bool)
    wvalid = np.ones(len(wv), bool) if wvalid is None else np.asarray(wvalid, bool)
    assert len(gv) == len(wv), f"{what}: {len(gv)} groups vs {len(wv)}"
    assert np.array_equal(gvalid, wvalid), f"{what}: validity differs"
    gv = np.asarray(gv)[wvalid]
    wv = np.asarray(wv)[wvalid]
    if gv.dtype.kind == "f":
        if rtol is None:
            rtol = 1e-5 if gv.dtype == np.float32 else 1e-12
        g64 = gv.astype(np.float64)
        w64 = wv.astype(np.float64)
        both_nan = np.isnan(g64) & np.isnan(w64)
        same_inf = np.isinf(w64) & (g64 == w64)
        with np.errstate(invalid="ignore"):  # inf - inf; such groups pass through same_inf
            err = np.abs(g64 - w64)
        ok = err <= rtol * np.abs(w64)
        if abs_sum is not None:
            ok |= err <= rtol * np.asarray(abs_sum)[wvalid]
        ok |= both_nan | same_inf
        assert ok.all(), (f"{what}: {int((~ok).sum())} of {ok.size} groups outside tolerance; "
                          f"worst err {err[~ok].max() if (~ok).any() else 0}")
    else:
        assert np.array_equal(gv, wv), f"{what}: integer results differ at {np.flatnonzero(gv != wv)[:5]}"
