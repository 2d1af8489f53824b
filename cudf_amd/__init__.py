"""cudf_amd: an MI355X-native hash groupby-aggregate with cudf's API.

The compute path is libcudf_amd.so (hand-written HIP kernels for gfx950 behind
the C ABI in include/mgb_groupby.h). This package is its Python binding,
shaped like pylibcudf's groupby."""
from . import aggregation, hashing
from .column import Column, ListColumn, Table, pack_validity, unpack_validity
from .batch import GroupByBatch
from .groupby import GroupBy, GroupByRequest
from .pool import DevicePool
from .serving import Server
from .types import Interpolation, Kind, NanEquality, NullEquality, NullPolicy, Sorted, TypeId

__all__ = [
    "aggregation", "hashing", "Column", "ListColumn", "Table", "GroupBy", "GroupByBatch", "GroupByRequest",
    "DevicePool", "Server", "Interpolation", "Kind", "NanEquality", "NullEquality", "NullPolicy", "Sorted", "TypeId",
    "pack_validity", "unpack_validity",
]
