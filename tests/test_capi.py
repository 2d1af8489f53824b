"""The C-ABI boundary on the CPU: libcudf_amd.so loads and exports every
entry point include/mgb_groupby.h declares, and argument errors are reported
before any device work (cpp/src/groupby/groupby.cu:208-216)."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

from cudf_amd import _lib
from cudf_amd.column import pack_validity, unpack_validity
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mgb_groupby.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mgb_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_functions()
    assert "mgb_groupby_aggregate" in names and len(names) >= 8
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "the Python binding must declare every C entry point"


def test_version():
    assert _lib.load().mgb_version() >= 100


@pytest.mark.parametrize("vt", [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21])
@pytest.mark.parametrize("kind", [0, 1, 2, 3, 4, 5, 8, 9, 10, 11, 12, 15, 16])
def test_target_type_matches_oracle(vt, kind):
    assert _lib.load().mgb_target_type(vt, kind) == O.target_type(vt, kind)


def _col(type_id, size, ptr=0x1000):
    c = _lib.mgb_column()
    c.type_id = type_id
    c.size = size
    c.data = ptr
    c.null_count = 0
    return c


def _call(keys, reqs, nout):
    lib = _lib.load()
    ka = (_lib.mgb_column * max(len(keys), 1))(*keys)
    ra = (_lib.mgb_request * max(len(reqs), 1))()
    keep = []
    for i, (vals, kinds) in enumerate(reqs):
        arr = (_lib.mgb_aggregation * len(kinds))(*[_lib.mgb_aggregation(k, 1) for k in kinds])
        keep.append(arr)
        ra[i].values = vals
        ra[i].aggs = ctypes.cast(arr, ctypes.POINTER(_lib.mgb_aggregation))
        ra[i].num_aggs = len(kinds)
    ok = (_lib.mgb_out_column * max(len(keys), 1))()
    orr = (_lib.mgb_out_column * max(nout, 1))()
    g = ctypes.c_int64(-1)
    calls = []
    alloc = _lib.mgb_allocator(_lib.ALLOC_FN(lambda c, n, s: calls.append(n) or None),
                               _lib.DEALLOC_FN(lambda c, p, n, s: None), None)
    st = lib.mgb_groupby_aggregate(ka, len(keys), 0, ra, len(reqs), ok, orr, ctypes.byref(g), None,
                                   ctypes.byref(alloc))
    return st, g.value, ok, orr, calls


def test_size_mismatch_is_a_logic_error():
    st, *_ = _call([_col(3, 10)], [(_col(3, 9), [0])], 1)
    assert st == _lib.MGB_ERROR_LOGIC
    assert b"Size mismatch" in _lib.load().mgb_last_error_message()


def test_invalid_type_aggregation_is_a_logic_error():
    st, *_ = _call([_col(3, 10)], [(_col(12, 10), [0])], 1)  # SUM of TIMESTAMP_DAYS
    assert st == _lib.MGB_ERROR_LOGIC
    assert b"Invalid type/aggregation" in _lib.load().mgb_last_error_message()
    with pytest.raises(_lib.LogicError):
        _lib.check(st)


def test_empty_input_returns_typed_empty_columns_without_allocating():
    st, g, ok, orr, calls = _call([_col(3, 0, 0)], [(_col(9, 0, 0), [0, 4, 9])], 3)
    assert st == _lib.MGB_SUCCESS and g == 0 and calls == []
    assert ok[0].type_id == 3 and ok[0].size == 0
    assert [orr[i].type_id for i in range(3)] == [9, 3, 10]  # SUM f32 -> f32, COUNT -> int32, MEAN -> f64


@pytest.mark.parametrize("rows", [2 ** 31, 2 ** 32, 2 ** 40])
def test_more_rows_than_size_type_refused_before_device_work(rows):
    # cudf::size_type is int32 (cpp/include/cudf/types.hpp:80): a call whose
    # columns hold 2^31 rows or more is refused, never run with 32-bit row
    # references that would wrap (the data pointer is never dereferenced)
    st, g, ok, orr, calls = _call([_col(3, rows)], [(_col(3, rows), [0, 4])], 2)
    assert st == _lib.MGB_ERROR_UNSUPPORTED and calls == []
    assert b"size_type" in _lib.load().mgb_last_error_message()
    lib = _lib.load()
    est = ctypes.c_int64(-1)
    key = (_lib.mgb_column * 1)(_col(4, rows))
    alloc = _lib.mgb_allocator(_lib.ALLOC_FN(lambda c, n, s: None), _lib.DEALLOC_FN(lambda c, p, n, s: None), None)
    assert lib.mgb_estimate_groups(key, 1, 0, None, ctypes.byref(alloc), ctypes.byref(est)) == \
        _lib.MGB_ERROR_UNSUPPORTED


@pytest.mark.parametrize("which", ["key", "value"])
def test_column_with_rows_and_no_data_is_a_logic_error(which):
    # the kernels read row 0 of a column in place of an absent row, so a
    # column with rows must have data (refused before any device work)
    key = _col(3, 10, 0 if which == "key" else 0x1000)
    val = _col(3, 10, 0 if which == "value" else 0x1000)
    st, g, ok, orr, calls = _call([key], [(val, [0])], 1)
    assert st == _lib.MGB_ERROR_LOGIC and calls == []
    assert b"has no data" in _lib.load().mgb_last_error_message()


def test_null_arguments_rejected():
    lib = _lib.load()
    g = ctypes.c_int64()
    st = lib.mgb_groupby_aggregate(None, 1, 0, None, 0, None, None, ctypes.byref(g), None, None)
    assert st == _lib.MGB_ERROR_LOGIC


@pytest.mark.parametrize("n", [0, 1, 31, 32, 33, 1000])
def test_validity_pack_roundtrip(n):
    rng = np.random.default_rng(n)
    v = rng.random(n) > 0.3
    words = pack_validity(v)
    assert (words.size * 4) % 64 == 0
    assert np.array_equal(unpack_validity(words, n), v)
    # LSB-first: bit i of word i//32
    for i in range(min(n, 40)):
        assert ((int(words[i // 32]) >> (i % 32)) & 1) == int(v[i])


def _no_device() -> bool:
    import torch

    return not torch.cuda.is_available()


HASH_KINDS = [0, 1, 2, 3, 4, 5, 8, 9, 10, 11, 12, 15, 16]  # every kind of the hash path (and M2)
SORT_KINDS = [13, 14, 17, 18]                             # MEDIAN, QUANTILE, NUNIQUE, NTH_ELEMENT


@pytest.mark.skipif(not _no_device(), reason="fake device pointers: run only where no GPU can read them")
@pytest.mark.parametrize("keys", [[3], [4], [9], [3, 4], [1, 10, 6]], ids=["i32", "i64", "f32", "i32_i64", "3keys"])
@pytest.mark.parametrize("vt", [1, 3, 4, 6, 9, 10, 11])
def test_every_request_is_planned_before_device_work(keys, vt):
    """On a machine without a GPU a valid call builds its plan (record
    layout, ops, outputs) and stops at the first device call; an invalid
    type/aggregation pair stops before it. The host sanitizer build runs
    this planning code (tests/test_sanitizers.py)."""
    lib = _lib.load()
    n = 1000
    kcols = []
    for t in keys:
        c = _col(t, n)
        c.null_mask = 0x2000
        c.null_count = 7
        kcols.append(c)
    vals = _col(vt, n)
    vals.null_mask = 0x3000
    vals.null_count = 5
    for kinds in (HASH_KINDS, SORT_KINDS, HASH_KINDS + SORT_KINDS):
        valid = [k for k in kinds if O.target_type(vt, k) != 0]
        for policy in (0, 1):
            ka = (_lib.mgb_column * len(kcols))(*kcols)
            ra = (_lib.mgb_request * 2)()
            aggs = (_lib.mgb_aggregation * len(valid))(*[_lib.mgb_aggregation(k, 1, 0, 0, 0.5) for k in valid])
            bad = (_lib.mgb_aggregation * 1)(_lib.mgb_aggregation(kinds[0], 1, 0, 0, 0.5))
            ra[0].values = vals
            ra[0].aggs = ctypes.cast(aggs, ctypes.POINTER(_lib.mgb_aggregation))
            ra[0].num_aggs = len(valid)
            ra[1].values = vals
            ra[1].aggs = ctypes.cast(bad, ctypes.POINTER(_lib.mgb_aggregation))
            ra[1].num_aggs = 1
            ok = (_lib.mgb_out_column * len(kcols))()
            orr = (_lib.mgb_out_column * (len(valid) + 1))()
            g = ctypes.c_int64(-1)
            alloc = _lib.mgb_allocator(_lib.ALLOC_FN(lambda c, b, s: None), _lib.DEALLOC_FN(lambda c, p, b, s: None),
                                       None)
            st = lib.mgb_groupby_aggregate(ka, len(kcols), policy, ra, 1, ok, orr, ctypes.byref(g), None,
                                           ctypes.byref(alloc))
            assert st in (_lib.MGB_ERROR_HIP, _lib.MGB_ERROR_ALLOC), lib.mgb_last_error_message()
            if O.target_type(vt, kinds[0]) == 0:
                st = lib.mgb_groupby_aggregate(ka, len(kcols), policy, ra, 2, ok, orr, ctypes.byref(g), None,
                                               ctypes.byref(alloc))
                assert st == _lib.MGB_ERROR_LOGIC
