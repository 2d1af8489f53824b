"""Serving over device-resident tables (``mgb_server_*`` in
include/mgb_groupby.h): queries submitted from any thread are batched
continuously by the library's scheduler thread, admitted against an HBM
budget, and run with overlapping device work. No counterpart in the
reference.

    srv = Server(hbm_budget_bytes=200 << 30)
    t = srv.add_table([key_col, f64_col, i32_col])
    ticket = srv.submit(t, [0], [(1, [agg.sum(), agg.mean()])])
    groups, latency_ms = srv.wait(ticket)
    keys, results = srv.take(ticket)
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence, Tuple

import torch

from . import _lib
from .aggregation import Aggregation
from .column import Column, Table
from .groupby import _TorchAllocator, _to_mgb, wrap_out_column
from .types import NullPolicy


class Server:
    def __init__(self, hbm_budget_bytes: int = 200 << 30, max_inflight: int = 2, max_batch_queries: int = 256,
                 batching: bool = True, reserve_bytes: int = 0, device: Optional[torch.device] = None):
        self._lib = _lib.load()
        opt = _lib.mgb_server_options()
        opt.hbm_budget_bytes = int(hbm_budget_bytes)
        opt.reserve_bytes = int(reserve_bytes)
        opt.max_inflight = int(max_inflight)
        opt.max_batch_queries = int(max_batch_queries)
        opt.batching = 1 if batching else 0
        h = ctypes.c_void_p()
        _lib.check(self._lib.mgb_server_create(ctypes.byref(opt), ctypes.byref(h)))
        self._h = h
        self._tables: List[List[Column]] = []
        self._aggs_per_ticket = {}
        self.device = device or torch.device("cuda", torch.cuda.current_device())

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mgb_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def add_table(self, columns: Sequence[Column]) -> int:
        cols = list(columns)
        arr = (_lib.mgb_column * len(cols))(*[_to_mgb(c) for c in cols])
        tid = ctypes.c_int32()
        _lib.check(self._lib.mgb_server_add_table(self._h, arr, len(cols), ctypes.byref(tid)))
        self._tables.append(cols)  # the server reads them until it is closed
        return tid.value

    def submit(self, table: int, key_columns: Sequence[int],
               requests: Sequence[Tuple[int, Sequence[Aggregation]]],
               null_handling: NullPolicy = NullPolicy.EXCLUDE) -> int:
        keys = (ctypes.c_int32 * len(key_columns))(*key_columns)
        reqs = (_lib.mgb_server_request * max(len(requests), 1))()
        keep = []
        for i, (col, aggs) in enumerate(requests):
            arr = (_lib.mgb_aggregation * max(len(aggs), 1))(
                *[a.to_c() for a in aggs])
            keep.append(arr)
            reqs[i].column = int(col)
            reqs[i].num_aggs = len(aggs)
            reqs[i].aggs = ctypes.cast(arr, ctypes.POINTER(_lib.mgb_aggregation))
        ticket = ctypes.c_int64()
        _lib.check(self._lib.mgb_server_submit(self._h, table, keys, len(key_columns), int(NullPolicy(null_handling)),
                                               reqs, len(requests), ctypes.byref(ticket)))
        self._aggs_per_ticket[ticket.value] = (len(key_columns), [len(a) for _, a in requests])
        return ticket.value

    def wait(self, ticket: int, timeout_ms: float = -1.0) -> Tuple[int, float]:
        g, lat = ctypes.c_int64(), ctypes.c_double()
        _lib.check(self._lib.mgb_server_wait(self._h, ticket, float(timeout_ms), ctypes.byref(g), ctypes.byref(lat)))
        return g.value, lat.value

    def take(self, ticket: int, stream: Optional[torch.cuda.Stream] = None):
        """(unique keys Table, [Table per request]) of a finished query; releases the ticket."""
        nk, per_req = self._aggs_per_ticket.pop(ticket)
        s = stream or torch.cuda.current_stream()
        out_keys = (_lib.mgb_out_column * max(nk, 1))()
        out_res = (_lib.mgb_out_column * max(sum(per_req), 1))()
        g = ctypes.c_int64()
        alloc = _TorchAllocator(self.device)
        _lib.check(self._lib.mgb_server_take_result(self._h, ticket, out_keys, out_res, ctypes.byref(g),
                                                    ctypes.c_void_p(s.cuda_stream), ctypes.byref(alloc.struct)))

        def wrap(o):
            return wrap_out_column(alloc, o)

        keys = Table([wrap(out_keys[i]) for i in range(nk)])
        results, j = [], 0
        for n in per_req:
            results.append(Table([wrap(out_res[j + i]) for i in range(n)]))
            j += n
        return keys, results

    def release(self, ticket: int) -> None:
        self._aggs_per_ticket.pop(ticket, None)
        _lib.check(self._lib.mgb_server_release(self._h, ticket))

    def stats(self) -> dict:
        st = _lib.mgb_server_stats()
        _lib.check(self._lib.mgb_server_get_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}
