"""Batched groupby-aggregate over the C ABI (``mgb_batch_*`` in
include/mgb_groupby.h): many independent queries with one schema planned
together and run as one HIP graph. No counterpart in the reference, which runs
one ``groupby.aggregate`` per call (python/cudf/cudf/_lib/pylibcudf/groupby.pyx:125-155).

    batch = GroupByBatch([(keys_table, [GroupByRequest(values, aggs)]), ...])
    batch.run()                    # execute + finish
    keys, results = batch.result(q)
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .column import Column, Table
from .groupby import GroupByRequest, _TorchAllocator, _to_mgb
from .types import NullPolicy

MODE_DIRECT, MODE_GLOBAL, MODE_PARTITIONED = 0, 1, 2
MODE_NAMES = {MODE_DIRECT: "lds_direct", MODE_GLOBAL: "global", MODE_PARTITIONED: "partitioned"}


class GroupByBatch:
    """Plans a batch of groupby queries that share one schema.

    ``queries``: (keys Table, [GroupByRequest]) per query; request r of every
    query must carry the same aggregations and value type. The input columns
    must stay alive and unchanged while the batch exists (the batch keeps
    references to them).
    """

    def __init__(self, queries: Sequence[Tuple[Table, List[GroupByRequest]]],
                 null_handling: NullPolicy = NullPolicy.EXCLUDE, force_mode: Optional[int] = None,
                 use_graph: bool = True, direct_item_rows: int = 0, estimate_scale: float = 1.0,
                 estimates: Optional[Sequence[int]] = None, stream: Optional[torch.cuda.Stream] = None):
        self._lib = _lib.load()
        if not queries:
            raise ValueError("a batch needs at least one query")
        self.queries = list(queries)
        nk = self.queries[0][0].num_columns()
        nreq = len(self.queries[0][1])
        self.num_keys, self.num_requests = nk, nreq
        self.aggs_per_request = [len(r.aggregations) for r in self.queries[0][1]]
        Q = len(self.queries)
        keys = (_lib.mgb_column * (Q * nk))()
        reqs = (_lib.mgb_request * max(Q * nreq, 1))()
        self._keep = []
        for q, (kt, rq) in enumerate(self.queries):
            if kt.num_columns() != nk or len(rq) != nreq:
                raise ValueError("every query of a batch needs the same number of key columns and requests")
            for c in range(nk):
                keys[q * nk + c] = _to_mgb(kt.columns[c])
            for r, req in enumerate(rq):
                arr = (_lib.mgb_aggregation * max(len(req.aggregations), 1))(
                    *[a.to_c() for a in req.aggregations])
                self._keep.append(arr)
                reqs[q * nreq + r].values = _to_mgb(req.values)
                reqs[q * nreq + r].aggs = ctypes.cast(arr, ctypes.POINTER(_lib.mgb_aggregation))
                reqs[q * nreq + r].num_aggs = len(req.aggregations)
        opt = _lib.mgb_batch_options()
        opt.force_mode = -1 if force_mode is None else int(force_mode)
        opt.use_graph = 1 if use_graph else 0
        opt.direct_item_rows = int(direct_item_rows)
        opt.estimate_scale = float(estimate_scale)
        if estimates is not None:
            if len(estimates) != Q:
                raise ValueError("one estimate per query")
            self._estimates = (ctypes.c_int64 * Q)(*[int(e) for e in estimates])
            opt.estimates = ctypes.cast(self._estimates, ctypes.POINTER(ctypes.c_int64))
        self.stream = stream if stream is not None else torch.cuda.current_stream()
        handle = ctypes.c_void_p()
        _lib.check(self._lib.mgb_batch_create(keys, nk, int(NullPolicy(null_handling)), reqs, nreq, Q,
                                              ctypes.byref(opt), ctypes.c_void_p(self.stream.cuda_stream),
                                              ctypes.byref(handle)))
        self._handle = handle
        self.device = None
        for kt, _ in self.queries:
            for c in kt.columns:
                if c.data is not None:
                    self.device = c.data.device
                    break
            if self.device is not None:
                break
        if self.device is None:
            self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        h = getattr(self, "_handle", None)
        if h is not None and h.value:
            self._lib.mgb_batch_destroy(h)
            self._handle = None

    def execute(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        s = stream or self.stream
        _lib.check(self._lib.mgb_batch_execute(self._handle, ctypes.c_void_p(s.cuda_stream)))

    def finish(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        s = stream or self.stream
        _lib.check(self._lib.mgb_batch_finish(self._handle, ctypes.c_void_p(s.cuda_stream)))

    def run(self, stream: Optional[torch.cuda.Stream] = None) -> None:
        self.execute(stream)
        self.finish(stream)

    def num_groups(self, q: int) -> int:
        g = ctypes.c_int64()
        _lib.check(self._lib.mgb_batch_result(self._handle, q, None, None, ctypes.byref(g)))
        return g.value

    def mode(self, q: int) -> Tuple[str, int]:
        m, e = ctypes.c_int32(), ctypes.c_int64()
        _lib.check(self._lib.mgb_batch_query_mode(self._handle, q, ctypes.byref(m), ctypes.byref(e)))
        return MODE_NAMES[m.value], e.value

    def stats(self) -> dict:
        st = _lib.mgb_batch_stats()
        _lib.check(self._lib.mgb_batch_get_stats(self._handle, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def result(self, q: int, stream: Optional[torch.cuda.Stream] = None):
        """(unique keys Table, [Table of results per request]) of query q, copied into new columns."""
        s = stream or self.stream
        nk, nres = self.num_keys, sum(self.aggs_per_request)
        out_keys = (_lib.mgb_out_column * max(nk, 1))()
        out_res = (_lib.mgb_out_column * max(nres, 1))()
        g = ctypes.c_int64()
        alloc = _TorchAllocator(self.device)
        _lib.check(self._lib.mgb_batch_copy_result(self._handle, q, out_keys, out_res, ctypes.byref(g),
                                                   ctypes.c_void_p(s.cuda_stream), ctypes.byref(alloc.struct)))

        def wrap(o):
            data = alloc.take(o.data)
            mask_t = alloc.take(o.null_mask)
            return Column(o.type_id, o.size, data, mask_t.view(torch.int32) if mask_t is not None else None,
                          o.null_count)

        keys = Table([wrap(out_keys[i]) for i in range(nk)])
        results, j = [], 0
        for n in self.aggs_per_request:
            results.append(Table([wrap(out_res[j + i]) for i in range(n)]))
            j += n
        return keys, results
