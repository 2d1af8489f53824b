"""Row hashing entry points (cudf::hashing::murmurhash3_x86_32 and
spark_murmurhash3_x86_32, cpp/include/cudf/hashing.hpp:90-134)."""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .column import Column, Table
from .groupby import _to_mgb
from .types import TypeId


def _hash(fn_name: str, table: Table, seed: int, out_type: TypeId) -> Column:
    lib = _lib.load()
    n = table.num_rows()
    dev = table.columns[0].data.device if table.columns and table.columns[0].data is not None else "cuda"
    out = torch.empty(max(n, 1) * 4, dtype=torch.uint8, device=dev)
    cols = (_lib.mgb_column * max(table.num_columns(), 1))(*[_to_mgb(c) for c in table.columns])
    stream = torch.cuda.current_stream()
    st = getattr(lib, fn_name)(cols, table.num_columns(), ctypes.c_uint32(seed & 0xFFFFFFFF),
                               ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream.cuda_stream))
    _lib.check(st)
    return Column(out_type, n, out)


def murmurhash3_x86_32(table: Table, seed: int = 0) -> Column:
    return _hash("mgb_murmurhash3_x86_32", table, seed, TypeId.UINT32)


def spark_murmurhash3_x86_32(table: Table, seed: int = 0) -> Column:
    return _hash("mgb_spark_murmurhash3_x86_32", table, seed, TypeId.INT32)
