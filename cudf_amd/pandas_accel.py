"""pandas groupby-aggregate on the MI355X path.

The reference runs pandas code on the GPU with cudf.pandas: every pandas call
is tried on cudf first and handed to pandas when cudf cannot serve it
(python/cudf/cudf/pandas/fast_slow_proxy.py:870-903). This module does that
for the one operation this library implements, ``DataFrame.groupby(by).agg``:

* ``groupby_agg(df, by, func, sort=True, dropna=True)`` runs the aggregation
  through ``cudf_amd.GroupBy`` (libcudf_amd.so) and returns what pandas
  returns for it, or raises ``NotServed`` for a call the library does not
  serve (a key or value dtype it has no kernel for, an aggregation it does
  not implement, keyword arguments, ...).
* ``install()`` patches ``agg`` / ``aggregate`` and the argument-free forms
  of the methods in ``METHODS`` (``sum()``, ``median()``, ``nunique()``, ...)
  on ``DataFrameGroupBy`` and ``SeriesGroupBy``, so that served calls take
  that path and every other call runs in pandas as before (calls pandas
  makes inside such a call included); ``stats()`` counts the calls each side
  took.

pandas' conventions, applied to the library's cudf-typed results (as cudf's
own Python layer does, python/cudf/cudf/core/groupby/groupby.py:627-655):
NaN values are skipped (they enter as nulls) and NaN keys are missing keys;
``sum`` / ``prod`` of a group without values is 0 / 1 and ``nunique`` 0,
other aggregations give NaN (or ``<NA>`` for nullable dtypes); ``sum``,
``prod``, ``min``, ``max``, ``first`` and ``last`` keep the column's dtype,
``count``, ``size`` and ``nunique`` are int64, ``mean``, ``median``, ``var``
and ``std`` are float64 except for float32 columns; ``idxmin`` / ``idxmax``
return index labels; ``first`` / ``last`` are the first / last value that is
not missing, ``nunique`` counts distinct values that are not missing; ``sort=True`` orders the groups by key,
``sort=False`` by first appearance.

Nothing here computes an aggregate on the host: when the library is not
built, ``groupby_agg`` raises and ``install()`` refuses.
"""
from __future__ import annotations

import threading
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import aggregation as agg
from .types import Kind, NullPolicy, TypeId

# pandas aggregation name -> (library kind, ddof)
AGGS = {
    "sum": (Kind.SUM, 0), "prod": (Kind.PRODUCT, 0), "min": (Kind.MIN, 0), "max": (Kind.MAX, 0),
    "count": (Kind.COUNT_VALID, 0), "size": (Kind.COUNT_ALL, 0), "mean": (Kind.MEAN, 0),
    "var": (Kind.VARIANCE, 1), "std": (Kind.STD, 1), "idxmin": (Kind.ARGMIN, 0), "idxmax": (Kind.ARGMAX, 0),
    # the sort-path kinds (missing values excluded, as pandas' defaults)
    "median": (Kind.MEDIAN, 0), "nunique": (Kind.NUNIQUE, 0), "first": (Kind.NTH_ELEMENT, 0),
    "last": (Kind.NTH_ELEMENT, -1),
}

_NUMPY_TYPES = {
    np.dtype(np.int8): TypeId.INT8, np.dtype(np.int16): TypeId.INT16, np.dtype(np.int32): TypeId.INT32,
    np.dtype(np.int64): TypeId.INT64, np.dtype(np.uint8): TypeId.UINT8, np.dtype(np.uint16): TypeId.UINT16,
    np.dtype(np.uint32): TypeId.UINT32, np.dtype(np.uint64): TypeId.UINT64, np.dtype(np.float32): TypeId.FLOAT32,
    np.dtype(np.float64): TypeId.FLOAT64, np.dtype(np.bool_): TypeId.BOOL8,
}


class NotServed(Exception):
    """The call is outside what the library serves; pandas runs it."""


def _column(series) -> Tuple[np.ndarray, np.ndarray, TypeId, object]:
    """(values, validity or None, library type, pandas dtype) of a column:
    numpy numeric dtypes (NaN is missing) and pandas' nullable numeric
    dtypes (``<NA>`` is missing)."""
    import pandas as pd

    dt = series.dtype
    if isinstance(dt, pd.api.extensions.ExtensionDtype):
        np_dt = getattr(dt, "numpy_dtype", None)
        if np_dt is None or np.dtype(np_dt) not in _NUMPY_TYPES:
            raise NotServed(f"dtype {dt}")
        arr = series.array
        valid = ~np.asarray(arr.isna())
        values = arr.to_numpy(dtype=np_dt, na_value=np.zeros((), np_dt))
    else:
        if np.dtype(dt) not in _NUMPY_TYPES:
            raise NotServed(f"dtype {dt}")
        values = series.to_numpy()
        valid = ~np.isnan(values) if values.dtype.kind == "f" else None
    if valid is not None and valid.all():
        valid = None
    return np.ascontiguousarray(values), valid, _NUMPY_TYPES[np.dtype(values.dtype)], dt


class Plan:
    """What one groupby-agg call asks for, in the library's terms."""

    def __init__(self, df, by, func, sort: bool, dropna: bool):
        import pandas as pd

        self.by = [by] if not isinstance(by, (list, tuple)) else list(by)
        if not self.by or any(not isinstance(b, (str, int)) or b not in df.columns for b in self.by):
            raise NotServed("keys other than column labels")
        if len(set(self.by)) != len(self.by) or len(self.by) > 8:
            raise NotServed("key columns")
        self.sort, self.dropna = bool(sort), bool(dropna)
        self.index = df.index
        self.rows = len(df)
        if self.rows == 0:
            raise NotServed("empty frame")  # pandas' empty results take dtypes from the spec's form
        self.keys = [_column(df[b]) for b in self.by]
        value_cols = [c for c in df.columns if c not in self.by]
        # (output label, column label, aggregation name)
        self.outputs: List[Tuple[object, object, str]] = []
        if isinstance(func, str) or (isinstance(func, (list, tuple)) and func):
            names = [func] if isinstance(func, str) else list(func)
            multi = not isinstance(func, str)
            for c in value_cols:
                for a in names:
                    self.outputs.append(((c, a) if multi else c, c, a))
        elif isinstance(func, dict) and func:
            multi = any(not isinstance(v, str) for v in func.values())
            for c, v in func.items():
                if c not in df.columns or c in self.by:
                    raise NotServed("aggregated column")
                if not isinstance(v, (str, list, tuple)):
                    raise NotServed(f"aggregation {v!r}")
                names = [v] if isinstance(v, str) else list(v)
                for a in names:
                    self.outputs.append(((c, a) if multi else c, c, a))
        else:
            raise NotServed("aggregation spec")
        self.multi = multi
        for _, _, a in self.outputs:
            if not isinstance(a, str) or a not in AGGS:
                raise NotServed(f"aggregation {a!r}")
        if not self.outputs:
            raise NotServed("nothing to aggregate")
        if len({lbl for lbl, _, _ in self.outputs}) != len(self.outputs):
            raise NotServed("duplicate outputs")
        # one request per value column, its aggregations in output order
        self.columns: Dict[object, Tuple] = {}
        self.requests: List[Tuple[object, List[str]]] = []
        for _, c, a in self.outputs:
            if c not in self.columns:
                self.columns[c] = _column(df[c])
                self.requests.append((c, []))
            req = next(r for r in self.requests if r[0] == c)
            if a not in req[1]:
                req[1].append(a)
        for c, names in self.requests:
            tid = self.columns[c][2]
            if tid == TypeId.BOOL8 and any(a in ("mean", "median", "var", "std", "sum", "prod") for a in names):
                raise NotServed("arithmetic on bool")
        self.pd = pd


def _device_inputs(plan: Plan, device: str):
    from .column import Column, Table
    from .groupby import GroupByRequest

    keys = Table([Column.from_numpy(v, valid, tid, device=device) for v, valid, tid, _ in plan.keys])
    reqs = []
    for c, names in plan.requests:
        v, valid, tid, _ = plan.columns[c]
        aggs = [agg.Aggregation(AGGS[a][0], AGGS[a][1]) for a in names]
        reqs.append(GroupByRequest(Column.from_numpy(v, valid, tid, device=device), aggs))
    if not plan.sort:
        # the first row of each group orders the groups as pandas does
        import torch

        first = Column.from_torch(torch.arange(plan.rows, device=device, dtype=torch.int64), TypeId.INT64)
        reqs.append(GroupByRequest(first, [agg.min()]))
    return keys, reqs


def _masked(pd, values: np.ndarray, ok: np.ndarray, dtype):
    """A nullable pandas array of `values` with <NA> where not `ok`."""
    arr = pd.array(values, dtype=dtype)
    if not ok.all():
        arr[~ok] = pd.NA
    return arr


def _as_pandas(plan: Plan, name: str, dtype, values: np.ndarray, valid):
    """One result column in pandas' conventions (see the module docstring)."""
    pd = plan.pd
    ok = np.ones(len(values), bool) if valid is None else np.asarray(valid, bool)
    nullable = isinstance(dtype, pd.api.extensions.ExtensionDtype)
    np_dt = np.dtype(getattr(dtype, "numpy_dtype", dtype))
    if name == "nunique":
        return values.astype(np.int64)  # int64 for nullable columns too
    if name in ("count", "size"):
        out = values.astype(np.int64)
        return pd.array(out, dtype="Int64") if nullable else out
    if name in ("idxmin", "idxmax"):
        labels = plan.index.take(np.where(ok, values, 0).astype(np.int64))
        if ok.all():
            return np.asarray(labels)
        return np.where(ok, np.asarray(labels, dtype=np.float64), np.nan)
    if name in ("sum", "prod"):
        empty = 0 if name == "sum" else 1
        if np_dt.kind in "iu":
            # pandas accumulates integers in 64 bits and casts the result back
            # to the column's dtype only when every group's value survives the
            # cast (its maybe_downcast_to_dtype); otherwise it keeps int64 /
            # uint64 (Int64 / UInt64 for nullable columns)
            wide = np.dtype(np.uint64 if np_dt.kind == "u" else np.int64)
            out = np.where(ok, values, empty).astype(wide)
            if np_dt.itemsize < 8 and np.array_equal(out.astype(np_dt).astype(wide), out):
                out, out_dt = out.astype(np_dt), dtype
            else:
                out_dt = ("UInt64" if np_dt.kind == "u" else "Int64") if nullable else wide
            return pd.array(out, dtype=out_dt) if nullable else out
        out = np.where(ok, values, empty).astype(np_dt) if np_dt.kind != "b" else values
        return pd.array(out, dtype=dtype) if nullable else out
    if name in ("min", "max", "first", "last"):
        if nullable:
            return _masked(pd, np.where(ok, values, 0).astype(np_dt), ok, dtype)
        if np_dt.kind == "f":
            return np.where(ok, values, np.nan).astype(np_dt)
        return values.astype(np_dt)  # numpy ints: every group has a value
    # mean, median, var, std: float64 (float32 for float32 columns)
    out_dt = np.float32 if np_dt == np.float32 else np.float64
    out = np.where(ok, values, np.nan).astype(out_dt)
    if nullable:
        return _masked(pd, np.where(ok, out, 0), ok, "Float32" if out_dt == np.float32 else "Float64")
    return out


def groupby_agg(df, by, func, sort: bool = True, dropna: bool = True, device: str = "cuda"):
    """``df.groupby(by, sort=sort, dropna=dropna).agg(func)`` on the MI355X
    path; raises NotServed for calls outside the library's scope."""
    plan = Plan(df, by, func, sort, dropna)
    from .groupby import GroupBy

    keys, reqs = _device_inputs(plan, device)
    gb = GroupBy(keys, NullPolicy.EXCLUDE if plan.dropna else NullPolicy.INCLUDE)
    out_keys, results = gb.aggregate(reqs)
    key_cols = [c.to_numpy() for c in out_keys.columns]
    res = {}
    for (c, names), table in zip(plan.requests, results):
        for a, col in zip(names, table.columns):
            res[(c, a)] = col.to_numpy()
    first = results[-1].columns[0].to_numpy()[0] if not plan.sort else None
    return _frame(plan, key_cols, res, first)


def _frame(plan: Plan, key_cols, res, first=None):
    """The pandas result from the library's key columns and results
    ({(column, aggregation): (values, valid)}); `first`: each group's first
    row (sort=False)."""
    pd = plan.pd
    n = len(key_cols[0][0]) if key_cols else 0
    levels = []
    for (kv, kvalid), (_, _, _, kdt) in zip(key_cols, plan.keys):
        ok = np.ones(n, bool) if kvalid is None else np.asarray(kvalid, bool)
        if isinstance(kdt, pd.api.extensions.ExtensionDtype):
            levels.append(_masked(pd, kv, ok, kdt))
        elif np.dtype(kdt).kind == "f":
            levels.append(np.where(ok, kv, np.nan).astype(kdt))
        else:
            levels.append(kv.astype(kdt))
    if len(plan.by) == 1:
        index = pd.Index(levels[0], name=plan.by[0])
    else:
        index = pd.MultiIndex.from_arrays(levels, names=plan.by)
    data = {}
    for label, c, a in plan.outputs:
        v, valid = res[(c, a)]
        data[label] = _as_pandas(plan, a, plan.columns[c][3], v, valid)
    out = pd.DataFrame(data, index=index)
    if plan.multi:
        out.columns = pd.MultiIndex.from_tuples(list(data.keys()))
    if plan.sort:
        out = out.sort_index(na_position="last")
    else:
        out = out.iloc[np.argsort(first, kind="stable")]
    return out


_lock = threading.Lock()
_counts = {"gpu": 0, "pandas": 0}
_saved = {}


def stats() -> Dict[str, int]:
    """Calls that took the MI355X path ("gpu") and calls left to pandas."""
    with _lock:
        return dict(_counts)


def _served_frame(gb, func):
    """The frame to aggregate when the library serves ``gb.agg(func)``, else None."""
    if not gb.as_index or gb.level is not None or func is None:
        return None
    try:
        df = gb.obj
        sel = getattr(gb, "_selection", None)
        if sel is not None:
            cols = [sel] if not isinstance(sel, (list, tuple)) else list(sel)
            keys = [gb.keys] if not isinstance(gb.keys, (list, tuple)) else list(gb.keys)
            df = df[keys + [c for c in cols if c not in keys]]
        Plan(df, gb.keys, func, gb.sort, gb.dropna)
        return df
    except Exception:  # NotServed, or a grouping this module does not read
        return None


def _series_frame(gb, func):
    """For ``SeriesGroupBy.agg(func)`` (``df.groupby(keys)[name]``) the
    library serves: the frame of the key columns and the series, else None."""
    if not gb.as_index or gb.level is not None or func is None or isinstance(func, dict):
        return None
    try:
        import pandas as pd

        groupings = gb._grouper.groupings
        keys = [gb.keys] if not isinstance(gb.keys, (list, tuple)) else list(gb.keys)
        name = gb.obj.name
        if ([g.name for g in groupings] != keys or not all(getattr(g, "in_axis", False) for g in groupings)
                or name is None or name in keys):
            return None
        data = {k: g.grouping_vector for k, g in zip(keys, groupings)}
        data[name] = gb.obj.array
        df = pd.DataFrame(data, index=gb.obj.index)
        Plan(df, keys, func, gb.sort, gb.dropna)
        return df
    except Exception:  # NotServed, or a grouping this module does not read
        return None


def _run_series(gb, df, func):
    """A SeriesGroupBy result: the series' column of the frame's result (a
    Series for one aggregation name, a frame of the names for a list)."""
    return _run(gb, df, func)[gb.obj.name]


def _run(gb, df, func):
    out = groupby_agg(df, gb.keys, func, gb.sort, gb.dropna)
    with _lock:
        _counts["gpu"] += 1
    return out


def _patched(self, func=None, *args, engine=None, engine_kwargs=None, **kwargs):
    df = None
    if not args and not kwargs and engine is None and engine_kwargs is None and not getattr(_tls, "inside", False):
        df = _served_frame(self, func)
    if df is None:
        return _in_pandas(_saved["aggregate"], self, func, *args, engine=engine, engine_kwargs=engine_kwargs, **kwargs)
    return _run(self, df, func)


# DataFrameGroupBy methods served when called without arguments (the same
# result as ``agg(name)``); with any argument pandas runs them
METHODS = ("sum", "prod", "min", "max", "count", "mean", "median", "var", "std", "nunique", "first", "last",
           "idxmin", "idxmax")


_tls = threading.local()  # .inside: within a call that pandas runs


def _in_pandas(orig, *args, **kwargs):
    """Runs a call in pandas; the patched methods that pandas calls inside it
    run in pandas too and are not counted."""
    if getattr(_tls, "inside", False):
        return orig(*args, **kwargs)
    with _lock:
        _counts["pandas"] += 1
    _tls.inside = True
    try:
        return orig(*args, **kwargs)
    finally:
        _tls.inside = False


def _series_patched(self, func=None, *args, engine=None, engine_kwargs=None, **kwargs):
    df = None
    if not args and not kwargs and engine is None and engine_kwargs is None and not getattr(_tls, "inside", False):
        df = _series_frame(self, func)
    if df is None:
        return _in_pandas(_saved["series.aggregate"], self, func, *args, engine=engine, engine_kwargs=engine_kwargs,
                          **kwargs)
    return _run_series(self, df, func)


def _method(name, series=False):
    key = ("series." if series else "") + name

    def patched(self, *args, **kwargs):
        df = None
        if not (args or kwargs or getattr(_tls, "inside", False)):
            df = _series_frame(self, name) if series else _served_frame(self, name)
        if df is None:
            return _in_pandas(_saved[key], self, *args, **kwargs)
        return _run_series(self, df, name) if series else _run(self, df, name)

    patched.__name__ = name
    return patched


def install() -> None:
    """Routes served ``DataFrameGroupBy.agg`` / ``SeriesGroupBy.agg`` calls,
    and the argument-free calls of the methods in METHODS on either, to the
    MI355X path. Raises when the library is not built."""
    from . import _lib
    from pandas.core.groupby.generic import DataFrameGroupBy, SeriesGroupBy

    _lib.load()
    with _lock:
        if "aggregate" in _saved:
            return
        _saved["aggregate"] = DataFrameGroupBy.aggregate
        _saved["agg"] = DataFrameGroupBy.agg
        DataFrameGroupBy.aggregate = _patched
        DataFrameGroupBy.agg = _patched
        _saved["series.aggregate"] = SeriesGroupBy.aggregate
        _saved["series.agg"] = SeriesGroupBy.agg
        SeriesGroupBy.aggregate = _series_patched
        SeriesGroupBy.agg = _series_patched
        for name in METHODS:
            _saved[name] = getattr(DataFrameGroupBy, name)
            setattr(DataFrameGroupBy, name, _method(name))
            _saved["series." + name] = getattr(SeriesGroupBy, name)
            setattr(SeriesGroupBy, name, _method(name, series=True))


def uninstall() -> None:
    from pandas.core.groupby.generic import DataFrameGroupBy, SeriesGroupBy

    with _lock:
        if "aggregate" not in _saved:
            return
        DataFrameGroupBy.aggregate = _saved.pop("aggregate")
        DataFrameGroupBy.agg = _saved.pop("agg")
        SeriesGroupBy.aggregate = _saved.pop("series.aggregate")
        SeriesGroupBy.agg = _saved.pop("series.agg")
        for name in METHODS:
            setattr(DataFrameGroupBy, name, _saved.pop(name))
            setattr(SeriesGroupBy, name, _saved.pop("series." + name))


__all__: Sequence[str] = ["groupby_agg", "install", "uninstall", "stats", "NotServed", "Plan"]
