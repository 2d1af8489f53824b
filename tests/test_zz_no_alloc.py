"""The allocation-free entry point (mgb_groupby_aggregate_noalloc) through
the C ABI: its results equal mgb_groupby_aggregate's on the same inputs (keys
and integers exact, floats within the stated tolerance), on two calls in a
row (the second smaller, over the scratch memory the first left behind), it
takes no device memory, and it refuses inputs beyond its limits."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

import cudf_amd as ca
from cudf_amd import _lib
from cudf_amd import aggregation as agg
from cudf_amd.groupby import _to_mgb
from cudf_amd.types import type_size
from tests.helpers import assert_same, sorted_result

pytestmark = pytest.mark.gpu

AGGS = [agg.sum(), agg.count(), agg.min(), agg.max(), agg.mean(), agg.variance(), agg.argmax()]


def _bitmask_bytes(n):
    return ((n + 31) // 32 * 4 + 63) // 64 * 64


def _inputs(rng, rows, groups, key_nulls=False):
    pool = rng.integers(-2 ** 62, 2 ** 62, groups, dtype=np.int64)
    k = pool[rng.integers(0, groups, rows)]
    kvalid = rng.random(rows) > 0.05 if key_nulls else None
    cols = [ca.Column.from_numpy(k, kvalid, ca.TypeId.INT64)]
    vals = [ca.Column.from_numpy(rng.integers(-1_000_000, 1_000_000, rows).astype(np.int32), rng.random(rows) > 0.05,
                                 ca.TypeId.INT32),
            ca.Column.from_numpy(rng.standard_normal(rows), rng.random(rows) > 0.05, ca.TypeId.FLOAT64)]
    return cols, vals


@pytest.fixture(scope="module", autouse=True)
def _warm(gpu):
    """One small call first: the runtime loads the library's code objects
    (and with them the scratch array) at the first launch on a device, which
    the measured calls must not count."""
    rng = np.random.default_rng(70)
    keys, vals = _inputs(rng, 1000, 10)
    st, _ = _noalloc(keys, vals, [agg.sum()], max_groups=10, check_memory=False)
    assert st == _lib.MGB_SUCCESS


def _noalloc(keys, vals, aggs, max_groups, null_policy=0, check_memory=True):
    """Calls the entry point with output buffers of max_groups rows; returns
    (status, (keys, results) as numpy or None)."""
    lib = _lib.load()
    nk = len(keys)
    ka = (_lib.mgb_column * nk)(*[_to_mgb(c) for c in keys])
    ra = (_lib.mgb_request * len(vals))()
    keep, out_types = [], []
    for i, v in enumerate(vals):
        arr = (_lib.mgb_aggregation * len(aggs))(*[a.to_c() for a in aggs])
        keep.append(arr)
        ra[i].values = _to_mgb(v)
        ra[i].aggs = ctypes.cast(arr, ctypes.POINTER(_lib.mgb_aggregation))
        ra[i].num_aggs = len(aggs)
        out_types += [lib.mgb_target_type(int(v.type_id), int(a.kind)) for a in aggs]
    bufs = []

    def col(o, type_id):
        data = torch.empty(max(max_groups * type_size(type_id), 1), dtype=torch.uint8, device="cuda")
        mask = torch.empty(_bitmask_bytes(max(max_groups, 1)) // 4, dtype=torch.int32, device="cuda")
        bufs.append((data, mask))
        o.data, o.data_bytes = data.data_ptr(), data.numel()
        o.null_mask, o.mask_bytes = mask.data_ptr(), mask.numel() * 4

    ok = (_lib.mgb_out_column * nk)()
    orr = (_lib.mgb_out_column * len(out_types))()
    for c in range(nk):
        col(ok[c], int(keys[c].type_id))
    for j, t in enumerate(out_types):
        col(orr[j], t)
    g = ctypes.c_int64(-1)
    stream = torch.cuda.current_stream()
    torch.cuda.synchronize()
    before = torch.cuda.mem_get_info()[0]
    st = lib.mgb_groupby_aggregate_noalloc(ka, nk, null_policy, ra, len(vals), ok, orr, max_groups, ctypes.byref(g),
                                           ctypes.c_void_p(stream.cuda_stream))
    after = torch.cuda.mem_get_info()[0]
    if check_memory:
        assert after == before, f"the call took {before - after} bytes of device memory"
    if st != _lib.MGB_SUCCESS:
        return st, None

    def host(o, i):
        data, mask = bufs[i]
        m = mask if o.null_mask else None
        assert o.size == g.value
        return ca.Column(o.type_id, o.size, data, m, o.null_count).to_numpy()

    kk = [host(ok[c], c) for c in range(nk)]
    res, j = [], 0
    for _ in vals:
        res.append([host(orr[j + a], nk + j + a) for a in range(len(aggs))])
        j += len(aggs)
    return st, (kk, res)


def _plain(keys, vals, aggs, null_policy=0):
    k, r = ca.GroupBy(ca.Table(keys), ca.NullPolicy(null_policy)).aggregate([ca.GroupByRequest(v, aggs) for v in vals])
    return [c.to_numpy() for c in k.columns], [[c.to_numpy() for c in t.columns] for t in r]


def _abs_sums(keys, vals, null_policy):
    """Per group, sorted by key: the sum of |x| of every float column (the
    scale of the float tolerance for groups whose sum cancels)."""
    out = []
    for v in vals:
        if v.type_id != ca.TypeId.FLOAT64:
            out.append(None)
            continue
        av = ca.Column.from_torch(v.data.view(torch.float64).abs(), ca.TypeId.FLOAT64, v.mask, v.null_count)
        k, r = ca.GroupBy(ca.Table(keys), ca.NullPolicy(null_policy)).aggregate([ca.GroupByRequest(av, [agg.sum()])])
        sk, sr = sorted_result([c.to_numpy() for c in k.columns], [[r[0].columns[0].to_numpy()]])
        out.append(sr[0][0][0])
    return out


def _same(got, want, abs_sums):
    gk, gr = sorted_result(*got)
    wk, wr = sorted_result(*want)
    assert len(gk[0][0]) == len(wk[0][0])
    for (gv, gvalid), (wv, wvalid) in zip(gk, wk):
        assert_same((gv, gvalid), (wv, wvalid), None, what="keys")
    for r, (gb, wb) in enumerate(zip(gr, wr)):
        count = wb[1][0]
        for a, (g, w) in enumerate(zip(gb, wb)):
            # integers, counts, MIN / MAX and ARGMAX exact; float SUM within
            # 1e-12 of the group's sum of |x|, MEAN of its mean of |x|,
            # VARIANCE within 1e-9 relative
            scale = None
            if abs_sums[r] is not None and a in (0, 4):
                scale = abs_sums[r] if a == 0 else abs_sums[r] / np.maximum(count, 1)
            assert_same(g, w, None, abs_sum=scale, rtol=1e-9 if a == 5 else None, what=f"request {r} agg {AGGS[a]}")


@pytest.mark.parametrize("null_policy", [0, 1])
def test_noalloc_matches_the_allocating_entry_point_twice(gpu, null_policy):
    rng = np.random.default_rng(71 + null_policy)
    # first call: partitioned execution (200K groups); second, smaller: the
    # direct execution over what the first left in the scratch memory
    for rows, groups in ((1_500_000, 150_000), (300_000, 700)):
        keys, vals = _inputs(rng, rows, groups, key_nulls=null_policy == 1)
        st, got = _noalloc(keys, vals, AGGS, max_groups=groups + 1, null_policy=null_policy)
        assert st == _lib.MGB_SUCCESS, _lib.load().mgb_last_error_message()
        _same(got, _plain(keys, vals, AGGS, null_policy), _abs_sums(keys, vals, null_policy))


def test_noalloc_limits(gpu):
    rng = np.random.default_rng(72)
    keys, vals = _inputs(rng, 100_000, 5000)
    # fewer output rows than groups: refused after the count, no write past the buffers
    st, _ = _noalloc(keys, vals, [agg.sum(), agg.count()], max_groups=4000)
    assert st == _lib.MGB_ERROR_ALLOC
    assert b"output buffers" in _lib.load().mgb_last_error_message()
    # and the next call is unaffected
    st, got = _noalloc(keys, vals, [agg.sum(), agg.count()], max_groups=5000)
    assert st == _lib.MGB_SUCCESS
    # more rows than the entry point takes: refused before any device work
    n_big = (1 << 23) + 1
    big = ca.Column(ca.TypeId.INT64, n_big, torch.empty(n_big * 8, dtype=torch.uint8, device=gpu))
    st, _ = _noalloc([big], [], [], max_groups=10)
    assert st == _lib.MGB_ERROR_UNSUPPORTED
    assert b"noalloc" in _lib.load().mgb_last_error_message()
