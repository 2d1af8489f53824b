"""Host-side entry points of the oracle (TEST INFRASTRUCTURE ONLY).

Two checkers live here:

* ``groupby`` / ``murmurhash3_table`` / ``spark_murmurhash3_table`` call the
  C restatement of libcudf's hash groupby and row hashers
  (oracle/groupby_oracle.c, built into oracle/liboracle.so by its Makefile).
* ``pandas_groupby`` runs the reference's own CPU path for this operation:
  the cudf.pandas fallback executes plain pandas
  (python/cudf/cudf/pandas/module_accelerator.py:79-92), so pandas'
  groupby-agg is called directly and its output is mapped to cudf's C++
  result types and null rules (Appendix: all-null SUM is null, integer SUM is
  int64, COUNT is int32, MEAN is float64).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
module. The product library never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from typing import List, Optional, Sequence, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# MGB_ORACLE_LIB: another build of the same source (the sanitizer build)
LIB_PATH = os.environ.get("MGB_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

# cudf::type_id values and their numpy storage types
INT8, INT16, INT32, INT64, UINT8, UINT16, UINT32, UINT64, FLOAT32, FLOAT64, BOOL8 = range(1, 12)
NP = {
    1: np.int8, 2: np.int16, 3: np.int32, 4: np.int64, 5: np.uint8, 6: np.uint16, 7: np.uint32,
    8: np.uint64, 9: np.float32, 10: np.float64, 11: np.uint8, 12: np.int32, 13: np.int64, 14: np.int64,
    15: np.int64, 16: np.int64, 17: np.int32, 18: np.int64, 19: np.int64, 20: np.int64, 21: np.int64,
}
# cudf::aggregation::Kind values
SUM, PRODUCT, MIN, MAX, COUNT_VALID, COUNT_ALL = 0, 1, 2, 3, 4, 5
SUM_OF_SQUARES, MEAN, M2, VARIANCE, STD, ARGMAX, ARGMIN = 8, 9, 10, 11, 12, 15, 16
MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT = 13, 14, 17, 18
COLLECT_LIST, COLLECT_SET = 21, 22
LIST = 24


class _Col(ctypes.Structure):
    _fields_ = [("type_id", ctypes.c_int32), ("data", ctypes.c_void_p), ("mask", ctypes.c_void_p)]


_lib = None


def build() -> str:
    """Compiles liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.mgo_murmur3_x86_32.restype = ctypes.c_uint32
        lib.mgo_murmur3_x86_32.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32]
        lib.mgo_spark_murmur3_x86_32.restype = ctypes.c_uint32
        lib.mgo_spark_murmur3_x86_32.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint32]
        lib.mgo_hash_combine.restype = ctypes.c_uint32
        lib.mgo_hash_combine.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        lib.mgo_target_type.restype = ctypes.c_int
        lib.mgo_target_type.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.mgo_murmurhash3_x86_32_table.restype = ctypes.c_int
        lib.mgo_murmurhash3_x86_32_table.argtypes = [ctypes.POINTER(_Col), ctypes.c_int32, ctypes.c_int64,
                                                     ctypes.c_uint32, ctypes.c_void_p]
        lib.mgo_spark_murmurhash3_x86_32_table.restype = ctypes.c_int
        lib.mgo_spark_murmurhash3_x86_32_table.argtypes = lib.mgo_murmurhash3_x86_32_table.argtypes
        lib.mgo_groupby_aggregate_ex.restype = ctypes.c_int
        lib.mgo_groupby_aggregate_ex.argtypes = [
            ctypes.POINTER(_Col), ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
            ctypes.POINTER(_Col), ctypes.c_int32,
            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32),
            ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_double),
            ctypes.c_int32, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
            ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p),
        ]
        lib.mgo_groupby_collect.restype = ctypes.c_int
        lib.mgo_groupby_collect.argtypes = [
            ctypes.POINTER(_Col), ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(_Col),
            ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
            ctypes.c_void_p, ctypes.c_void_p,
        ]
        _lib = lib
    return _lib


def pack_mask(valid: Optional[np.ndarray]) -> Optional[np.ndarray]:
    """bool[n] -> uint32 words, LSB first (cudf bitmask layout)."""
    if valid is None:
        return None
    valid = np.asarray(valid, dtype=bool)
    nwords = (valid.size + 31) // 32 + 1
    raw = np.zeros(nwords * 4, dtype=np.uint8)
    bits = np.packbits(valid, bitorder="little")
    raw[: bits.size] = bits
    return raw.view(np.uint32)


class HostColumn:
    """A host column for the oracle: storage array of the cudf type plus validity."""

    def __init__(self, type_id: int, values, valid: Optional[np.ndarray] = None):
        self.type_id = int(type_id)
        self.values = np.ascontiguousarray(np.asarray(values).astype(NP[self.type_id], copy=False))
        self.valid = None if valid is None else np.asarray(valid, dtype=bool)
        self._mask = pack_mask(self.valid)

    def c(self) -> _Col:
        return _Col(self.type_id, self.values.ctypes.data if self.values.size else None,
                    self._mask.ctypes.data if self._mask is not None else None)

    def __len__(self) -> int:
        return self.values.size


def murmur3(data: bytes, seed: int = 0) -> int:
    buf = ctypes.create_string_buffer(data, len(data))
    return load().mgo_murmur3_x86_32(buf, len(data), seed & 0xFFFFFFFF)


def spark_murmur3(data: bytes, seed: int = 0) -> int:
    buf = ctypes.create_string_buffer(data, len(data))
    return load().mgo_spark_murmur3_x86_32(buf, len(data), seed & 0xFFFFFFFF)


def hash_combine(lhs: int, rhs: int) -> int:
    return load().mgo_hash_combine(lhs & 0xFFFFFFFF, rhs & 0xFFFFFFFF)


def target_type(type_id: int, kind: int) -> int:
    return load().mgo_target_type(type_id, kind)


def murmurhash3_table(cols: Sequence[HostColumn], seed: int = 0) -> np.ndarray:
    n = len(cols[0])
    out = np.zeros(max(n, 1), dtype=np.uint32)
    arr = (_Col * len(cols))(*[c.c() for c in cols])
    rc = load().mgo_murmurhash3_x86_32_table(arr, len(cols), n, seed & 0xFFFFFFFF, out.ctypes.data)
    assert rc == 0
    return out[:n]


def spark_murmurhash3_table(cols: Sequence[HostColumn], seed: int = 0) -> np.ndarray:
    n = len(cols[0])
    out = np.zeros(max(n, 1), dtype=np.int32)
    arr = (_Col * len(cols))(*[c.c() for c in cols])
    rc = load().mgo_spark_murmurhash3_x86_32_table(arr, len(cols), n, seed & 0xFFFFFFFF, out.ctypes.data)
    assert rc == 0
    return out[:n]


def groupby(keys: Sequence[HostColumn], requests: Sequence[Tuple[HostColumn, Sequence]],
            include_null_keys: bool = False):
    """The C restatement of cudf's hash groupby.

    requests: [(values, [kind, (kind, ddof) or (kind, ddof, null_policy,
    interpolation, quantile), ...]), ...]; ddof is n for NTH_ELEMENT
    Returns (key columns [(values, valid)], results [[(values, valid), ...] per request]),
    groups in order of first appearance."""
    lib = load()
    n = len(keys[0]) if keys else 0
    vals: List[HostColumn] = []
    req_col, req_kind, req_ddof, req_nulls, req_interp, req_q, shape = [], [], [], [], [], [], []
    for v, aggs in requests:
        vals.append(v)
        shape.append(len(aggs))
        for a in aggs:
            spec = (a, 1) if isinstance(a, int) else tuple(a)
            kind, ddof = spec[0], spec[1]
            nulls, interp, q = spec[2:5] if len(spec) >= 5 else (0, 0, 0.5)
            req_col.append(len(vals) - 1)
            req_kind.append(kind)
            req_ddof.append(ddof)
            req_nulls.append(nulls)
            req_interp.append(interp)
            req_q.append(q)
    nreq = len(req_kind)
    out_data, out_valid = [], []
    for j in range(nreq):
        tt = target_type(vals[req_col[j]].type_id, req_kind[j])
        if tt < 0:
            raise ValueError("Invalid type/aggregation combination.")
        out_data.append(np.zeros(max(n, 1), dtype=NP[tt]))
        out_valid.append(np.zeros(max(n, 1), dtype=np.uint8))
    reps = np.zeros(max(n, 1), dtype=np.int64)
    ng = ctypes.c_int64(0)
    karr = (_Col * max(len(keys), 1))(*[k.c() for k in keys])
    varr = (_Col * max(len(vals), 1))(*[v.c() for v in vals])
    i32 = ctypes.c_int32 * max(nreq, 1)
    dptr = (ctypes.c_void_p * max(nreq, 1))(*[d.ctypes.data for d in out_data])
    vptr = (ctypes.c_void_p * max(nreq, 1))(*[d.ctypes.data for d in out_valid])
    f64 = ctypes.c_double * max(nreq, 1)
    rc = lib.mgo_groupby_aggregate_ex(karr, len(keys), n, 1 if include_null_keys else 0, varr, len(vals),
                                   i32(*req_col), i32(*req_kind), i32(*req_ddof), i32(*req_nulls),
                                   i32(*req_interp), f64(*req_q), nreq,
                                   reps.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(ng),
                                   dptr, vptr)
    if rc != 0:
        raise ValueError("Invalid type/aggregation combination.")
    g = ng.value
    reps = reps[:g]
    out_keys = []
    for k in keys:
        kv = k.values[reps]
        kvalid = k.valid[reps] if k.valid is not None else None
        out_keys.append((kv, kvalid))
    results, j = [], 0
    for cnt in shape:
        block = []
        for _ in range(cnt):
            block.append((out_data[j][:g].copy(), out_valid[j][:g].astype(bool)))
            j += 1
        results.append(block)
    return out_keys, results


def collect(keys: Sequence[HostColumn], values: HostColumn, kind: int, null_policy: int = 1,
            nulls_equal: int = 0, nans_equal: int = 0, include_null_keys: bool = False):
    """The C restatement of cudf's groupby COLLECT_LIST / COLLECT_SET.

    null_policy 1 = INCLUDE, 0 = EXCLUDE; nulls_equal 0 = EQUAL, 1 = UNEQUAL;
    nans_equal 0 = ALL_EQUAL, 1 = UNEQUAL. Returns (key columns [(values,
    valid)], lists [(values, valid) per group]), groups in order of first
    appearance; a set's elements are in order of first appearance."""
    lib = load()
    n = len(keys[0]) if keys else 0
    reps = np.zeros(max(n, 1), dtype=np.int64)
    offsets = np.zeros(max(n, 1) + 1, dtype=np.int64)
    rows = np.zeros(max(n, 1), dtype=np.int64)
    ng = ctypes.c_int64(0)
    karr = (_Col * max(len(keys), 1))(*[k.c() for k in keys])
    v = values.c()
    rc = lib.mgo_groupby_collect(karr, len(keys), n, 1 if include_null_keys else 0, ctypes.byref(v), kind,
                                 null_policy, nulls_equal, nans_equal, reps.ctypes.data, ctypes.byref(ng),
                                 offsets.ctypes.data, rows.ctypes.data)
    if rc != 0:
        raise ValueError("Invalid type/aggregation combination.")
    g = ng.value
    reps = reps[:g]
    out_keys = [(k.values[reps], k.valid[reps] if k.valid is not None else None) for k in keys]
    valid = values.valid if values.valid is not None else np.ones(len(values), bool)
    lists = []
    for i in range(g):
        r = rows[offsets[i]:offsets[i + 1]]
        lists.append((values.values[r].copy(), valid[r].copy()))
    return out_keys, lists


def sort_list(vals: np.ndarray, valid: np.ndarray):
    """A list's elements ascending, NaN after every number, nulls last: the
    order cudf::lists::sort_lists(ASCENDING, null_order::AFTER) gives and the
    reference's COLLECT_SET tests compare in
    (cpp/tests/groupby/collect_set_tests.cpp:58-63)."""
    vals = np.asarray(vals)
    valid = np.asarray(valid, bool)
    v = vals[valid]
    if v.dtype.kind == "f":
        order = np.lexsort((v, np.isnan(v)))  # numbers ascending, then NaNs
    else:
        order = np.argsort(v, kind="stable")
    nn = int((~valid).sum())
    out = np.concatenate([v[order], np.zeros(nn, dtype=vals.dtype)])
    return out, np.concatenate([np.ones(v.size, bool), np.zeros(nn, bool)])


def sort_order(key_cols) -> np.ndarray:
    """Row order sorting groups by key (nulls first, then by value), the
    comparison protocol of cpp/tests/groupby/groupby_test_util.cpp:40-95."""
    sort_keys = []
    for vals, valid in reversed(list(key_cols)):
        v = np.asarray(vals)
        if v.dtype.kind == "f":
            v = np.where(np.isnan(v), np.inf, v)
        if valid is not None:  # a null's payload is unspecified: nulls tie on it
            v = np.where(np.asarray(valid, bool), v, np.zeros((), v.dtype))
        sort_keys.append(v)
        if valid is not None:
            sort_keys.append(np.asarray(valid, dtype=np.int8))
    if not sort_keys:
        return np.zeros(0, dtype=np.int64)
    return np.lexsort(sort_keys)


# ---------------------------------------------------------------- pandas

PANDAS_AGG = {SUM: "sum", MIN: "min", MAX: "max", COUNT_VALID: "count", MEAN: "mean", COUNT_ALL: "size"}


def pandas_groupby(key_cols, value_cols, aggs_per_col, dropna: bool = True):
    """pandas groupby-agg (the cudf.pandas CPU fallback) mapped to cudf's C++ results.

    key_cols: [(values, valid-or-None)], value_cols: [(values, valid-or-None)],
    aggs_per_col: [[kind, ...]] (SUM, MIN, MAX, COUNT_VALID, MEAN, COUNT_ALL).
    Returns (keys [(values, valid)], results [[(values, valid)]]) sorted by key.
    """
    import pandas as pd

    data = {}
    for i, (v, valid) in enumerate(key_cols):
        s = pd.Series(v)
        if valid is not None:
            s = s.astype(_nullable(v.dtype)).mask(~np.asarray(valid, dtype=bool))
        data[f"k{i}"] = s
    for i, (v, valid) in enumerate(value_cols):
        s = pd.Series(v)
        if v.dtype.kind in "iub":
            s = s.astype("Int64")  # cudf sums integers in int64
        elif valid is not None:
            s = s.astype(_nullable(v.dtype))
        if valid is not None:
            s = s.mask(~np.asarray(valid, dtype=bool))
        data[f"v{i}"] = s
    df = pd.DataFrame(data)
    kn = [f"k{i}" for i in range(len(key_cols))]
    gb = df.groupby(kn, dropna=dropna, sort=True)
    out_keys = None
    results = []
    for i, kinds in enumerate(aggs_per_col):
        names = [PANDAS_AGG[k] for k in kinds]
        res = gb[f"v{i}"].agg(names)
        if out_keys is None:
            idx = res.index
            if len(kn) == 1:
                arrays = [idx]
            else:
                arrays = [idx.get_level_values(j) for j in range(len(kn))]
            out_keys = []
            for arr, (v, _) in zip(arrays, key_cols):
                a = pd.array(arr)
                valid = ~np.asarray(pd.isna(a))
                vals = np.asarray(pd.Series(a).fillna(0).to_numpy(), dtype=v.dtype)
                out_keys.append((vals, valid))
        cnt = gb[f"v{i}"].count().to_numpy().astype(np.int64)
        block = []
        for k, name in zip(kinds, names):
            col = res[name]
            src = value_cols[i][0]
            if k in (COUNT_VALID, COUNT_ALL):
                block.append((col.to_numpy().astype(np.int32), np.ones(len(col), dtype=bool)))
                continue
            valid = cnt > 0  # cudf: null iff the group has no valid value
            if k == MEAN:
                vals = pd.Series(col).astype("float64").fillna(0).to_numpy()
            elif k == SUM and src.dtype.kind in "iub":
                vals = pd.Series(col).fillna(0).to_numpy().astype(np.int64)
            elif src.dtype.kind in "iub":
                vals = pd.Series(col).fillna(0).to_numpy().astype(src.dtype)
            else:
                vals = pd.Series(col).astype(src.dtype.name if src.dtype.kind == "f" else "float64")
                vals = vals.fillna(0).to_numpy().astype(src.dtype)
            block.append((vals, valid))
        results.append(block)
    return out_keys, results


def _nullable(dt: np.dtype) -> str:
    return {"i": "Int", "u": "UInt", "f": "Float", "b": "boolean"}[dt.kind] + (
        str(dt.itemsize * 8) if dt.kind != "b" else "")
