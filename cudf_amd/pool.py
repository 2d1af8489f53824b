"""Device memory pool for the temporaries of groupby calls (``mgb_pool_*`` in
include/mgb_groupby.h; the role of rmm's pool resource in front of cudf).

Every device temporary of a call (hash tables, partitioned rows, histograms,
staged results, sort space) is taken from a pool and returned to it,
stream-ordered, when the call ends; results come from torch's allocator. A
pool never holds more than its limit::

    pool = DevicePool(8 << 30)
    GroupBy(keys, temporaries=pool).aggregate(requests)

Without one, calls on a device share that device's default pool, whose
limit is the device's memory until ``set_default_limit`` lowers it.
"""
from __future__ import annotations

import ctypes
import threading
from typing import Dict, Optional

import torch

from . import _lib


class DevicePool:
    """A bounded device memory pool; ``allocator`` is its mgb_allocator."""

    def __init__(self, limit_bytes: int):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.mgb_pool_create(int(limit_bytes), ctypes.byref(h)))
        self._h = h
        self.allocator = _lib.mgb_allocator()
        _lib.check(self._lib.mgb_pool_allocator(self._h, ctypes.byref(self.allocator)))

    def stats(self) -> dict:
        st = _lib.mgb_pool_stats()
        _lib.check(self._lib.mgb_pool_get_stats(self._h, ctypes.byref(st)))
        return {name: getattr(st, name) for name, _ in st._fields_}

    def trim(self) -> None:
        """Frees the chunks no block uses (waits for the streams that used them)."""
        _lib.check(self._lib.mgb_pool_trim(self._h))

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.mgb_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


_defaults: Dict[int, DevicePool] = {}
_limits: Dict[int, int] = {}
_lock = threading.Lock()


def default_pool(device: Optional[torch.device] = None) -> DevicePool:
    """The pool of `device` that calls without a pool of their own share."""
    idx = torch.device(device).index if device is not None else torch.cuda.current_device()
    if idx is None:
        idx = torch.cuda.current_device()
    with _lock:
        pool = _defaults.get(idx)
        if pool is None:
            limit = _limits.get(idx) or torch.cuda.get_device_properties(idx).total_memory
            pool = _defaults[idx] = DevicePool(limit)
        return pool


def set_default_limit(limit_bytes: int, device: Optional[torch.device] = None) -> None:
    """Bounds the default pool of `device` (a new pool; the old one is freed
    once the calls that use it have finished)."""
    idx = torch.device(device).index if device is not None else torch.cuda.current_device()
    with _lock:
        _limits[idx] = int(limit_bytes)
        old = _defaults.pop(idx, None)
    if old is not None:
        old.close()
