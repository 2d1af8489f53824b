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
