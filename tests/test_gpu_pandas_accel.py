"""The pandas hook on the GPU: ``df.groupby(by).agg(func)`` through
libcudf_amd.so against pandas itself (the reference's CPU path for this
call), for the frames of tests/test_pandas_accel.py and larger ones, and the
installed hook routing served calls to the library."""
from __future__ import annotations

import numpy as np
import pandas as pd
import pytest

from cudf_amd import pandas_accel as PA
from tests.test_pandas_accel import ALL, frame, same

pytestmark = pytest.mark.gpu


def test_list_of_aggregations(gpu):
    df = frame(1)
    sub = df[["k", "f64", "f32", "i8", "ni"]]
    same(PA.groupby_agg(sub, "k", ALL), sub.groupby("k").agg(ALL))


def test_dict_spec_idx_and_two_keys(gpu):
    df = frame(2).drop(columns="k2")
    spec = {"f64": ["mean", "idxmin", "idxmax", "var", "std"], "i8": "sum", "ni": ["min", "count"]}
    same(PA.groupby_agg(df, "k", spec), df.groupby("k").agg(spec))
    df5 = frame(5).drop(columns=["ni"])
    spec2 = {"f64": ["sum", "std"], "f32": ["min"]}
    same(PA.groupby_agg(df5, ["k", "k2"], spec2), df5.groupby(["k", "k2"]).agg(spec2))


def test_sort_false_and_dropna_false(gpu):
    df = frame(4, nan_keys=True).drop(columns=["k2", "ni"])
    spec = {"f64": ["sum", "mean"], "i8": ["max", "count"]}
    same(PA.groupby_agg(df, "k", spec, sort=False), df.groupby("k", sort=False).agg(spec))
    same(PA.groupby_agg(df, "k", spec, dropna=False), df.groupby("k", dropna=False).agg(spec))


@pytest.mark.parametrize("rows,card", [(2_000_000, 100_000), (1_000_000, 12)])
def test_large_frames(gpu, rows, card):
    rng = np.random.default_rng(rows + card)
    keys = rng.integers(-2 ** 40, 2 ** 40, card)[rng.integers(0, card, rows)]
    f64 = rng.standard_normal(rows)
    f64[rng.random(rows) < 0.05] = np.nan
    df = pd.DataFrame({"key": keys, "f64": f64, "i32": rng.integers(-1000, 1000, rows).astype(np.int32)})
    spec = {"f64": ["sum", "count", "mean", "min", "max", "std", "idxmax"], "i32": ["sum", "min", "var", "idxmin"]}
    same(PA.groupby_agg(df, "key", spec), df.groupby("key").agg(spec))


def test_sort_path_aggregations(gpu):
    df = frame(11)
    sub = df[["k", "f64", "f32", "i8", "ni"]]
    spec = ["median", "nunique", "first", "last"]
    same(PA.groupby_agg(sub, "k", spec), sub.groupby("k").agg(spec))
    spec2 = {"f64": ["median", "sum"], "ni": ["first", "nunique"], "i8": "last"}
    same(PA.groupby_agg(df, ["k", "k2"], spec2), df.groupby(["k", "k2"]).agg(spec2))
    same(PA.groupby_agg(sub, "k", spec, sort=False), sub.groupby("k", sort=False).agg(spec))
    rng = np.random.default_rng(12)
    n = 1_000_000
    big = pd.DataFrame({"key": rng.integers(0, 50_000, n), "x": np.round(rng.standard_normal(n), 2),
                        "y": rng.integers(-20, 20, n).astype(np.int32)})
    big.loc[rng.random(n) < 0.05, "x"] = np.nan
    spec3 = {"x": ["median", "nunique", "first", "last"], "y": ["median", "nunique"]}
    same(PA.groupby_agg(big, "key", spec3), big.groupby("key").agg(spec3))


def test_installed_hook_runs_on_the_library(gpu):
    df = frame(8)
    want = df.groupby("k").agg({"f64": ["sum", "var"], "i8": "max"})
    PA.install()
    try:
        before = PA.stats()
        got = df.groupby("k").agg({"f64": ["sum", "var"], "i8": "max"})
        sem = df.groupby("k").agg({"f64": "sem"})  # not served: pandas runs it
        after = PA.stats()
    finally:
        PA.uninstall()
    same(got, want)
    pd.testing.assert_frame_equal(sem, df.groupby("k").agg({"f64": "sem"}))
    assert after["gpu"] - before["gpu"] == 1 and after["pandas"] - before["pandas"] == 1


def test_installed_hook_serves_methods(gpu):
    df = frame(14)[["k", "f64", "f32", "i8", "ni"]]
    # products of the integer columns overflow here (pandas then widens the
    # dtype, the library wraps in it): prod runs on the float columns
    frames = {name: df[["k", "f64", "f32"]] if name == "prod" else df for name in PA.METHODS}
    want = {name: getattr(frames[name].groupby("k"), name)() for name in PA.METHODS}
    PA.install()
    try:
        before = PA.stats()
        got = {name: getattr(frames[name].groupby("k"), name)() for name in PA.METHODS}
        after = PA.stats()
    finally:
        PA.uninstall()
    for name in PA.METHODS:
        same(got[name], want[name])
    assert after["gpu"] - before["gpu"] == len(PA.METHODS)


def test_installed_hook_serves_series_groupby(gpu):
    df = frame(16)
    calls = [("f64", "sum"), ("f64", "median"), ("f32", "nunique"), ("i8", "first"), ("ni", "mean"),
             ("ni", "last"), ("f64", "idxmax")]
    want = [getattr(df.groupby("k")[c], m)() for c, m in calls]
    want_list = df.groupby(["k", "k2"])["f64"].agg(["sum", "median", "std"])
    PA.install()
    try:
        before = PA.stats()
        got = [getattr(df.groupby("k")[c], m)() for c, m in calls]
        got_list = df.groupby(["k", "k2"])["f64"].agg(["sum", "median", "std"])
        after = PA.stats()
    finally:
        PA.uninstall()
    for (c, m), g, w in zip(calls, got, want):
        pd.testing.assert_series_equal(g, w, check_exact=False, rtol=1e-5, atol=1e-4, check_dtype=False)
        assert g.dtype.kind == w.dtype.kind, (c, m, g.dtype, w.dtype)
    same(got_list, want_list)
    assert after["gpu"] - before["gpu"] == len(calls) + 1
