"""pylibcudf-shaped groupby over the C ABI (mirrors
python/cudf/cudf/_lib/pylibcudf/groupby.pyx:80-157): ``GroupBy(keys,
null_handling, keys_are_sorted).aggregate([GroupByRequest(values, aggs)])``
returns ``(unique_keys: Table, [Table of results per request])``."""
from __future__ import annotations

import ctypes
import threading
from typing import Dict, List, Sequence

import torch

from . import _lib
from .aggregation import Aggregation
from .column import Column, ListColumn, Table
from .pool import DevicePool, default_pool
from .types import NullPolicy, Sorted, TypeId


def _to_mgb(c: Column) -> _lib.mgb_column:
    m = _lib.mgb_column()
    m.type_id = int(c.type_id)
    m.size = c.size
    m.data = c.data_ptr or None
    m.null_mask = c.mask_ptr or None
    m.offset = c.offset
    m.null_count = c.null_count if c.mask is not None else 0
    return m


class _TorchAllocator:
    """mgb_allocator backed by torch's caching allocator: each buffer is a
    uint8 tensor kept alive here until the result columns take it over. All
    result buffers of one call come from one tensor (allocate_many): each
    column holds a view of it, and the tensor is freed with the last view."""

    def __init__(self, device: torch.device):
        self.device = device
        self.live: Dict[int, torch.Tensor] = {}
        self.lock = threading.Lock()
        self._alloc = _lib.ALLOC_FN(self._allocate)
        self._dealloc = _lib.DEALLOC_FN(self._deallocate)
        self._alloc_many = _lib.ALLOC_MANY_FN(self._allocate_many)
        self.struct = _lib.mgb_allocator(self._alloc, self._dealloc, None, self._alloc_many)

    def _allocate_many(self, ctx, n, sizes, out, stream):
        offs, total = [], 0
        for i in range(n):
            offs.append(total)
            total += (int(sizes[i]) + 255) // 256 * 256
        try:
            t = torch.empty(max(total, 1), dtype=torch.uint8, device=self.device)
        except Exception:  # out of memory: the library reports MGB_ERROR_ALLOC
            return 1
        base = t.data_ptr()
        with self.lock:
            for i in range(n):
                out[i] = base + offs[i]
                self.live[base + offs[i]] = t[offs[i]: offs[i] + int(sizes[i])]
        return 0

    def _allocate(self, ctx, nbytes, stream):
        try:
            t = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)
        except Exception:  # out of memory: the library reports MGB_ERROR_ALLOC
            return None
        with self.lock:
            self.live[t.data_ptr()] = t
        return t.data_ptr()

    def _deallocate(self, ctx, ptr, nbytes, stream):
        with self.lock:
            self.live.pop(ptr, None)

    def take(self, ptr: int):
        if not ptr:
            return None
        with self.lock:
            return self.live.pop(ptr)


def wrap_out_column(alloc, o: _lib.mgb_out_column):
    """Column (or ListColumn) over the device buffers of one result column
    that the library allocated through `alloc`."""
    data = alloc.take(o.data)
    mask_t = alloc.take(o.null_mask)
    mask = mask_t.view(torch.int32) if mask_t is not None else None
    if o.type_id == TypeId.LIST:
        offsets = Column(TypeId.INT32, o.size + 1 if o.size else 0, data)
        cmask_t = alloc.take(o.child_null_mask)
        child = Column(o.child_type_id, o.child_size, alloc.take(o.child_data),
                       cmask_t.view(torch.int32) if cmask_t is not None else None, o.child_null_count)
        return ListColumn(o.size, offsets, child)
    return Column(o.type_id, o.size, data, mask, o.null_count)


class GroupByRequest:
    """Values to aggregate and the aggregations to compute on them."""

    def __init__(self, values: Column, aggregations: Sequence[Aggregation]):
        self.values = values
        self.aggregations = list(aggregations)


class GroupBy:
    """Group values by keys and compute aggregations (cudf::groupby::groupby)."""

    def __init__(self, keys: Table, null_handling: NullPolicy = NullPolicy.EXCLUDE,
                 keys_are_sorted: Sorted = Sorted.NO, temporaries: "DevicePool" = None):
        self.keys = keys
        self.null_handling = NullPolicy(null_handling)
        self.keys_are_sorted = Sorted(keys_are_sorted)
        # the pool of every device temporary of a call (None: the device's default pool)
        self.temporaries = temporaries

    def aggregate(self, requests: List[GroupByRequest], stream: torch.cuda.Stream = None):
        lib = _lib.load()
        if stream is None:
            stream = torch.cuda.current_stream()
        dev = None
        for c in self.keys.columns:
            if c.data is not None:
                dev = c.data.device
                break
        if dev is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        nk = self.keys.num_columns()
        keys_arr = (_lib.mgb_column * max(nk, 1))(*[_to_mgb(c) for c in self.keys.columns])
        agg_arrays = []
        reqs = (_lib.mgb_request * max(len(requests), 1))()
        n_out = 0
        for i, r in enumerate(requests):
            arr = (_lib.mgb_aggregation * max(len(r.aggregations), 1))(
                *[a.to_c() for a in r.aggregations])
            agg_arrays.append(arr)
            reqs[i].values = _to_mgb(r.values)
            reqs[i].aggs = ctypes.cast(arr, ctypes.POINTER(_lib.mgb_aggregation))
            reqs[i].num_aggs = len(r.aggregations)
            n_out += len(r.aggregations)
        out_keys = (_lib.mgb_out_column * max(nk, 1))()
        out_res = (_lib.mgb_out_column * max(n_out, 1))()
        groups = ctypes.c_int64(0)
        alloc = _TorchAllocator(dev)
        pool = self.temporaries if self.temporaries is not None else default_pool(dev)
        alloc.struct.temporaries = ctypes.pointer(pool.allocator)
        st = lib.mgb_groupby_aggregate(keys_arr, nk, int(self.null_handling), reqs, len(requests),
                                       out_keys, out_res, ctypes.byref(groups),
                                       ctypes.c_void_p(stream.cuda_stream), ctypes.byref(alloc.struct))
        _lib.check(st)

        def wrap(o: _lib.mgb_out_column):
            return wrap_out_column(alloc, o)

        key_table = Table([wrap(out_keys[i]) for i in range(nk)])
        results = []
        j = 0
        for r in requests:
            cols = []
            for _ in r.aggregations:
                cols.append(wrap(out_res[j]))
                j += 1
            results.append(Table(cols))
        return key_table, results
