"""The pandas hook's host logic (cudf_amd.pandas_accel) on the CPU: which
calls it serves, and the conversion of cudf-typed results to pandas'
conventions, checked against pandas itself. The oracle stands in for the
device here (it returns the library's result shapes); the GPU tests run the
same frames through libcudf_amd.so."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from cudf_amd import pandas_accel as PA
from oracle import oracle as O


def frame(seed=0, n=2000, nan_keys=False):
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 40, n).astype(np.int32)
    k2 = rng.integers(0, 3, n).astype(np.int64)
    f64 = rng.standard_normal(n) * 10
    f64[rng.random(n) < 0.1] = np.nan
    f64[k == 7] = np.nan  # a group without values
    f32 = rng.standard_normal(n).astype(np.float32)
    i8 = rng.integers(-5, 6, n).astype(np.int8)
    ni = pd.array(rng.integers(-100, 100, n), dtype="Int32")
    ni[rng.random(n) < 0.2] = pd.NA
    df = pd.DataFrame({"k": k, "k2": k2, "f64": f64, "f32": f32, "i8": i8, "ni": ni})
    if nan_keys:
        kf = k.astype(np.float64)
        kf[rng.random(n) < 0.05] = np.nan
        df["k"] = kf
    return df


def oracle_frame(df, by, func, sort=True, dropna=True):
    """PA's result assembly with the oracle computing the aggregates."""
    plan = PA.Plan(df, by, func, sort, dropna)
    keys = [(int(tid), v, valid) for v, valid, tid, _ in plan.keys]
    reqs = []
    for c, names in plan.requests:
        v, valid, tid, _ = plan.columns[c]
        reqs.append((O.HostColumn(int(tid), v, valid), [(int(PA.AGGS[a][0]), PA.AGGS[a][1]) for a in names]))
    if not sort:
        reqs.append((O.HostColumn(O.INT64, np.arange(len(df), dtype=np.int64)), [O.MIN]))
    kcols = [O.HostColumn(t, v, valid) for t, v, valid in keys]
    key_cols, results = O.groupby(kcols, reqs, include_null_keys=not dropna)
    res = {}
    for (c, names), blk in zip(plan.requests, results):
        for a, col in zip(names, blk):
            res[(c, a)] = col
    first = results[-1][0][0] if not sort else None
    return PA._frame(plan, key_cols, res, first)


def same(got, want):
    # float32 sums accumulate at float32 precision in cudf (pandas sums them
    # in float64): groups that cancel near zero differ by ~1e-5 absolute.
    # dtypes: numpy columns exactly; pandas' nullable columns by kind and
    # nullability (pandas widens some of them depending on the spec's form)
    pd.testing.assert_frame_equal(got, want, check_exact=False, rtol=1e-5, atol=1e-4, check_dtype=False)
    for c in want.columns:
        g, w = got[c].dtype, want[c].dtype
        if isinstance(w, pd.api.extensions.ExtensionDtype):
            assert isinstance(g, pd.api.extensions.ExtensionDtype) and g.kind == w.kind, (c, g, w)
        else:
            assert g == w, (c, g, w)


ALL = ["sum", "count", "mean", "min", "max", "var", "std", "size"]


def test_list_of_aggregations():
    df = frame(1)
    sub = df[["k", "f64", "f32", "i8", "ni"]]
    same(oracle_frame(sub, "k", ALL), sub.groupby("k").agg(ALL))


def test_prod():
    # values whose products stay in range (an overflowing product wraps, and
    # pandas then widens some dtypes)
    rng = np.random.default_rng(9)
    n = 600
    df = pd.DataFrame({"k": rng.integers(0, 30, n).astype(np.int16),
                       "s": rng.choice(np.array([-1, 1], np.int32), n),
                       "f": rng.choice(np.array([0.5, 2.0, np.nan]), n)})
    same(oracle_frame(df, "k", ["prod", "sum"]), df.groupby("k").agg(["prod", "sum"]))
    same(oracle_frame(df, "k", {"s": "prod", "f": "prod"}), df.groupby("k").agg({"s": "prod", "f": "prod"}))


@pytest.mark.parametrize("dtype", ["int8", "int16", "int32", "uint8", "uint32", "Int8", "Int32", "UInt16"])
@pytest.mark.parametrize("func", ["sum", "prod"])
def test_narrow_integer_results_widen_like_pandas(dtype, func):
    """pandas keeps a narrow integer dtype for sum / prod only when every
    group's result fits it; otherwise int64 / uint64 (Int64 / UInt64)."""
    info = np.iinfo(pd.api.types.pandas_dtype(dtype).numpy_dtype if dtype[0] in "IU" else np.dtype(dtype))
    big = info.max // 2 + 1
    for vals in ([big, big, 1, 2], [1, 2, 1, 1]):
        df = pd.DataFrame({"k": np.array([1, 1, 2, 2], np.int32), "v": pd.array(vals, dtype=dtype)})
        got = oracle_frame(df, "k", [func])
        want = df.groupby("k").agg([func])
        pd.testing.assert_frame_equal(got, want, check_dtype=True)


def test_dict_spec_and_single_string():
    df = frame(2)
    spec = {"f64": ["mean", "idxmin", "idxmax"], "i8": "sum", "ni": ["min", "count"]}
    same(oracle_frame(df.drop(columns="k2"), "k", spec), df.drop(columns="k2").groupby("k").agg(spec))
    spec2 = {"f64": "max", "i8": "min"}
    same(oracle_frame(df, "k", spec2), df.groupby("k").agg(spec2))
    sub = df[["k", "f32", "i8"]]
    same(oracle_frame(sub, "k", "mean"), sub.groupby("k").agg("mean"))


def test_index_labels_for_idxmin():
    df = frame(3).drop(columns=["k2", "ni"])
    df.index = pd.Index(np.arange(len(df)) * 3 + 100)
    spec = {"f32": ["idxmin", "idxmax"], "i8": ["idxmax"]}
    same(oracle_frame(df, "k", spec), df.groupby("k").agg(spec))


def test_sort_false_and_dropna_false():
    df = frame(4, nan_keys=True).drop(columns=["k2", "ni"])
    spec = {"f64": ["sum", "mean"], "i8": ["max", "count"]}
    same(oracle_frame(df, "k", spec, sort=False), df.groupby("k", sort=False).agg(spec))
    same(oracle_frame(df, "k", spec, dropna=False), df.groupby("k", dropna=False).agg(spec))


def test_two_keys():
    df = frame(5).drop(columns=["ni"])
    spec = {"f64": ["sum", "std"], "f32": ["min"]}
    same(oracle_frame(df, ["k", "k2"], spec), df.groupby(["k", "k2"]).agg(spec))


def test_sort_path_aggregations():
    df = frame(11)
    sub = df[["k", "f64", "f32", "i8", "ni"]]
    spec = ["median", "nunique", "first", "last"]
    same(oracle_frame(sub, "k", spec), sub.groupby("k").agg(spec))
    spec2 = {"f64": ["median", "sum"], "ni": ["first", "nunique"], "i8": "last"}
    same(oracle_frame(df, ["k", "k2"], spec2), df.groupby(["k", "k2"]).agg(spec2))
    same(oracle_frame(sub, "k", spec, sort=False), sub.groupby("k", sort=False).agg(spec))


@pytest.mark.parametrize("func", ["sem", ["sum", "skew"], {"f64": "ohlc"}, {"f64": lambda x: x.sum()}, {}])
def test_calls_not_served(func):
    df = frame(6)
    with pytest.raises(PA.NotServed):
        PA.Plan(df, "k", func, True, True)


def test_empty_frame_not_served():
    with pytest.raises(PA.NotServed):
        PA.Plan(frame(1).iloc[:0], "k", "sum", True, True)


def test_dropna_false_without_missing_keys():
    df = frame(10).drop(columns=["k2", "ni"])
    spec = {"f64": ["sum", "max"], "i8": "count"}
    same(oracle_frame(df, "k", spec, dropna=False), df.groupby("k", dropna=False).agg(spec))


def test_dtypes_not_served():
    df = frame(7)
    df["s"] = ["x"] * len(df)
    with pytest.raises(PA.NotServed):
        PA.Plan(df, "k", "sum", True, True)  # a string column among the values
    with pytest.raises(PA.NotServed):
        PA.Plan(df, "s", {"f64": "sum"}, True, True)  # a string key


def test_install_routes_only_served_calls(monkeypatch):
    """With the hook installed, a call it does not serve still runs in pandas
    (and is counted there); served calls go to groupby_agg."""
    calls = []
    monkeypatch.setattr(PA, "groupby_agg", lambda df, by, func, sort, dropna: calls.append(func) or "gpu")
    PA.install()
    try:
        df = frame(8)
        before = PA.stats()
        want = df.groupby("k").agg({"f64": "sem"})
        got = df.groupby("k").agg({"f64": "sem"})
        pd.testing.assert_frame_equal(got, want)
        assert df.groupby("k").agg({"f64": "sum"}) == "gpu" and calls == [{"f64": "sum"}]
        assert df.groupby("k")[["f64", "i8"]].aggregate("max") == "gpu"
        after = PA.stats()
        assert after["pandas"] - before["pandas"] == 2 and after["gpu"] - before["gpu"] == 2
    finally:
        PA.uninstall()
    assert not isinstance(frame(8).groupby("k").agg({"f64": "sum"}), str)


def test_install_routes_argument_free_methods(monkeypatch):
    """``gb.sum()`` and the other methods in METHODS go to groupby_agg when
    called without arguments and give agg(name)'s result; with an argument
    pandas runs them."""
    calls = []
    monkeypatch.setattr(PA, "groupby_agg", lambda df, by, func, sort, dropna: calls.append(func) or "gpu")
    df = frame(8)
    want = df.groupby("k").sum(min_count=1)
    PA.install()
    try:
        for name in PA.METHODS:
            assert getattr(df.groupby("k"), name)() == "gpu"
        assert calls == list(PA.METHODS)
        pd.testing.assert_frame_equal(df.groupby("k").sum(min_count=1), want)
    finally:
        PA.uninstall()
    assert not isinstance(df.groupby("k").sum(), str)


@pytest.mark.parametrize("name", PA.METHODS)
def test_methods_equal_agg_by_name(name):
    """The method and agg(name) agree in pandas itself, so serving one as the
    other keeps pandas' result."""
    df = frame(13)[["k", "f64", "f32", "i8"]]
    pd.testing.assert_frame_equal(getattr(df.groupby("k"), name)(), df.groupby("k").agg(name))


def test_series_groupby_through_the_hook(monkeypatch):
    """``df.groupby(k)[x]`` calls: agg with a name or a list and the
    argument-free methods, with the oracle standing in for the device, equal
    pandas' results; a dict spec, an argument, or a key that is not a column
    stays in pandas."""
    monkeypatch.setattr(PA, "groupby_agg", lambda df, by, func, sort, dropna: oracle_frame(df, by, func, sort, dropna))
    df = frame(15)
    calls = [("f64", "agg", "sum"), ("f64", "agg", ["median", "nunique", "first"]), ("i8", "agg", ["mean", "max"])]
    calls += [(c, m, None) for c in ("f64", "ni") for m in PA.METHODS if m != "prod"]
    want = [(getattr(df.groupby("k")[c], m)(f) if f is not None else getattr(df.groupby("k")[c], m)())
            for c, m, f in calls]
    want_two = df.groupby(["k", "k2"])["f32"].median()
    want_unsorted = df.groupby("k", sort=False)["f64"].sum()
    PA.install()
    try:
        before = PA.stats()
        got = [(getattr(df.groupby("k")[c], m)(f) if f is not None else getattr(df.groupby("k")[c], m)())
               for c, m, f in calls]
        got_two = df.groupby(["k", "k2"])["f32"].median()
        got_unsorted = df.groupby("k", sort=False)["f64"].sum()
        mid = PA.stats()
        df.groupby("k")["f64"].sum(min_count=2)
        df.groupby(df["k"].to_numpy())["f64"].sum()
        after = PA.stats()
    finally:
        PA.uninstall()
    for (c, m, f), g, w in zip(calls, got, want):
        if isinstance(w, pd.Series):
            pd.testing.assert_series_equal(g, w, check_exact=False, rtol=1e-5, atol=1e-4, check_dtype=False)
            assert g.dtype == w.dtype or (g.dtype.kind == w.dtype.kind), (c, m, g.dtype, w.dtype)
        else:
            same(g, w)
    pd.testing.assert_series_equal(got_two, want_two, check_exact=False, rtol=1e-5)
    pd.testing.assert_series_equal(got_unsorted, want_unsorted, check_exact=False, rtol=1e-9)
    assert mid["gpu"] - before["gpu"] == len(calls) + 2
    assert after["pandas"] - mid["pandas"] >= 2 and after["gpu"] == mid["gpu"]
