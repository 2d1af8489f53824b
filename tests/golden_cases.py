"""Expands tests/golden/reference_groupby_cases.json (literal fixtures of the
reference's C++ groupby tests) into concrete typed cases."""
from __future__ import annotations

import json
import os

import numpy as np

from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
TYPES = {"INT8": O.INT8, "INT16": O.INT16, "INT32": O.INT32, "INT64": O.INT64, "UINT8": O.UINT8,
         "UINT16": O.UINT16, "UINT32": O.UINT32, "UINT64": O.UINT64, "FLOAT32": O.FLOAT32,
         "FLOAT64": O.FLOAT64, "BOOL8": O.BOOL8}
KINDS = {"SUM": O.SUM, "PRODUCT": O.PRODUCT, "MIN": O.MIN, "MAX": O.MAX, "COUNT_VALID": O.COUNT_VALID,
         "COUNT_ALL": O.COUNT_ALL, "SUM_OF_SQUARES": O.SUM_OF_SQUARES, "MEAN": O.MEAN,
         "VARIANCE": O.VARIANCE, "STD": O.STD, "ARGMAX": O.ARGMAX, "ARGMIN": O.ARGMIN, "M2": O.M2,
         "MEDIAN": O.MEDIAN, "QUANTILE": O.QUANTILE, "NUNIQUE": O.NUNIQUE, "NTH_ELEMENT": O.NTH_ELEMENT}
INTERP = {"LINEAR": 0, "LOWER": 1, "HIGHER": 2, "MIDPOINT": 3, "NEAREST": 4}
NULLS = {"EXCLUDE": 0, "INCLUDE": 1}


def agg_spec(c, kind):
    """The oracle's aggregation spec of a case: kind, (kind, ddof) or
    (kind, ddof or n, null_policy, interpolation, quantile)."""
    if kind in (O.MEDIAN, O.QUANTILE, O.NUNIQUE, O.NTH_ELEMENT):
        return (kind, c.get("n", 0), NULLS[c.get("nulls", "EXCLUDE")], INTERP[c.get("interp", "LINEAR")],
                c.get("q", 0.5))
    return (kind, c["ddof"]) if "ddof" in c else kind


def _num(x):
    if isinstance(x, str):
        return float(x)
    return x


def load_cases():
    with open(os.path.join(HERE, "golden", "reference_groupby_cases.json")) as f:
        doc = json.load(f)
    out = []
    for c in doc["cases"]:
        kind = KINDS[c["agg"]]
        if c.get("typed"):
            combos = [("INT32", v) for v in doc["value_types_for_typed_cases"]]
        elif c.get("typed_both"):  # keys and values of the same type (keys_col<T>, vals_col<T>)
            combos = [(v, v) for v in doc["value_types_for_typed_cases"]]
        elif c.get("typed_keys"):
            combos = [(k, c["val_type"]) for k in ("INT8", "INT16", "INT32", "INT64", "FLOAT32", "FLOAT64")]
        elif "val_types" in c:
            combos = [(c["key_type"], v) for v in c["val_types"]]
        else:
            combos = [(c["key_type"], c["val_type"])]
        for kt, vt in combos:
            k_t, v_t = TYPES[kt], TYPES[vt]
            keys = np.array([_num(x) for x in c["keys"]], dtype=O.NP[k_t])
            vals = np.array([_num(x) for x in c["vals"]], dtype=O.NP[v_t])
            kvalid = np.array(c["keys_valid"], bool) if "keys_valid" in c else None
            vvalid = np.array(c["vals_valid"], bool) if "vals_valid" in c else None
            tt = O.target_type(v_t, kind)
            exp_keys = np.array([_num(x) for x in c["expect_keys"]], dtype=O.NP[k_t])
            exp_keys_valid = np.array(c["expect_keys_valid"], bool) if "expect_keys_valid" in c else None
            exp_vals = np.array([_num(x) for x in c["expect_vals"]], dtype=np.float64).astype(O.NP[tt])
            exp_valid = np.array(c.get("expect_valid", [1] * len(c["expect_vals"])), bool)
            out.append(dict(
                id=f"{c['name']}[{kt}->{vt}]", cite=c["cite"], kind=agg_spec(c, kind),
                include_null_keys=c.get("null_policy") == "INCLUDE",
                keys=(k_t, keys, kvalid), vals=(v_t, vals, vvalid), target=tt,
                expect_keys=(exp_keys, exp_keys_valid), expect=(exp_vals, exp_valid)))
    return out


def check_case(case, key_cols, results):
    """Compares one groupby output (unsorted) with a case's expectation."""
    order = O.sort_order(key_cols)
    got_k, got_kv = key_cols[0]
    got_k = got_k[order]
    got_kv = None if got_kv is None else np.asarray(got_kv, bool)[order]
    exp_k, exp_kv = case["expect_keys"]
    eorder = O.sort_order([(exp_k, exp_kv)])
    exp_k = exp_k[eorder]
    exp_kv = None if exp_kv is None else exp_kv[eorder]
    assert len(got_k) == len(exp_k), f"{case['id']}: {len(got_k)} groups, expected {len(exp_k)}"
    gkv = np.ones(len(got_k), bool) if got_kv is None else got_kv
    ekv = np.ones(len(exp_k), bool) if exp_kv is None else exp_kv
    assert np.array_equal(gkv, ekv), f"{case['id']}: key validity"
    assert np.array_equal(got_k[ekv], exp_k[ekv]), f"{case['id']}: keys {got_k} vs {exp_k}"
    gv, gvalid = results[0][0]
    gv = np.asarray(gv)[order]
    gvalid = np.ones(len(gv), bool) if gvalid is None else np.asarray(gvalid, bool)[order]
    ev, evalid = case["expect"]
    ev, evalid = ev[eorder], evalid[eorder]
    assert np.array_equal(gvalid, evalid), f"{case['id']}: validity {gvalid} vs {evalid}"
    if ev.dtype.kind == "f":
        # CUDF_TEST_EXPECT_COLUMNS_EQUIVALENT compares floats to 4 ulp
        # (cpp/tests/utilities/column_utilities.cu:426)
        a, b = gv[evalid].astype(np.float64), ev[evalid].astype(np.float64)
        with np.errstate(invalid="ignore"):
            same = (a == b) | (np.abs(a - b) <= 4 * np.spacing(np.abs(b).astype(ev.dtype)).astype(np.float64))
            same |= np.isnan(a) & np.isnan(b)
        assert same.all(), f"{case['id']}: {gv} vs {ev}"
    else:
        assert np.array_equal(gv[evalid], ev[evalid]), f"{case['id']}: {gv} vs {ev}"


COLLECT_TYPES = dict(TYPES, TIMESTAMP_DAYS=12, TIMESTAMP_SECONDS=13, TIMESTAMP_NANOSECONDS=16)
COLLECT_KINDS = {"COLLECT_LIST": O.COLLECT_LIST, "COLLECT_SET": O.COLLECT_SET}
EQUALITY = {"EQUAL": 0, "UNEQUAL": 1, "ALL_EQUAL": 0}


def load_collect_cases():
    """tests/golden/reference_collect_cases.json expanded into typed cases:
    keys / values already sliced, the aggregation as (kind, null_policy,
    nulls_equal, nans_equal), expected lists as [(values, valid)] in the
    reference's sorted key order."""
    with open(os.path.join(HERE, "golden", "reference_collect_cases.json")) as f:
        doc = json.load(f)
    out = []
    for c in doc["cases"]:
        vts = doc["value_types_for_typed_cases"] if c.get("typed") else [c["val_type"]]
        lo, hi = c.get("slice", [0, len(c["keys"])])
        for vt in vts:
            v_t = COLLECT_TYPES[vt]
            dt = O.NP[v_t]
            keys = np.array(c["keys"], dtype=np.int32)[lo:hi]
            vals = np.array([_num(x) for x in c["vals"]], dtype=np.float64).astype(dt)[lo:hi]
            vvalid = np.array(c["vals_valid"], bool)[lo:hi] if "vals_valid" in c else None
            lists = []
            for lst in c["expect_lists"]:
                valid = np.array([x is not None for x in lst], bool)
                v = np.array([0 if x is None else _num(x) for x in lst], dtype=np.float64).astype(dt)
                lists.append((v, valid))
            out.append(dict(
                id=f"{c['name']}[{vt}]", cite=c["cite"],
                agg=(COLLECT_KINDS[c["agg"]], NULLS[c.get("nulls", "INCLUDE")],
                     EQUALITY[c.get("nulls_equal", "EQUAL")], EQUALITY[c.get("nans_equal", "ALL_EQUAL")]),
                keys=(O.INT32, keys, None), vals=(v_t, vals, vvalid),
                expect_keys=np.array(c["expect_keys"], dtype=np.int32), expect_lists=lists))
    return out


def same_list(got, want) -> bool:
    """Element-wise equality of two lists (values, valid): same validity, equal
    valid values (NaN equal to NaN, -0.0 to +0.0)."""
    gv, gvalid = np.asarray(got[0]), np.asarray(got[1], bool)
    wv, wvalid = np.asarray(want[0]), np.asarray(want[1], bool)
    if gv.size != wv.size or not np.array_equal(gvalid, wvalid):
        return False
    a, b = gv[wvalid], wv[wvalid]
    if a.dtype.kind == "f":
        return bool(((a == b) | (np.isnan(a) & np.isnan(b))).all())
    return bool(np.array_equal(a, b))


def check_collect_case(case, key_col, lists, sort_sets=True):
    """Compares one COLLECT output (groups unsorted) with a case: keys sorted,
    and each list as is (COLLECT_LIST, or COLLECT_SET with sort_sets False)
    or after sorting its elements (COLLECT_SET with sort_sets True)."""
    kv = np.asarray(key_col[0])
    order = np.argsort(kv, kind="stable")
    assert np.array_equal(kv[order], case["expect_keys"]), f"{case['id']}: keys {kv[order]}"
    for pos, g in enumerate(order):
        got = lists[g]
        if case["agg"][0] == O.COLLECT_SET and sort_sets:
            got = O.sort_list(*got)
        want = case["expect_lists"][pos]
        assert same_list(got, want), f"{case['id']}: group {kv[g]}: {got} vs {want}"
