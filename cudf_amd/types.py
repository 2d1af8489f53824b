"""Type ids, null policies and aggregation kinds numbered as in libcudf
(cpp/include/cudf/types.hpp:201-233, cpp/include/cudf/aggregation.hpp:85-122)."""
from __future__ import annotations

import enum

import numpy as np


class TypeId(enum.IntEnum):
    EMPTY = 0
    INT8 = 1
    INT16 = 2
    INT32 = 3
    INT64 = 4
    UINT8 = 5
    UINT16 = 6
    UINT32 = 7
    UINT64 = 8
    FLOAT32 = 9
    FLOAT64 = 10
    BOOL8 = 11
    TIMESTAMP_DAYS = 12
    TIMESTAMP_SECONDS = 13
    TIMESTAMP_MILLISECONDS = 14
    TIMESTAMP_MICROSECONDS = 15
    TIMESTAMP_NANOSECONDS = 16
    DURATION_DAYS = 17
    DURATION_SECONDS = 18
    DURATION_MILLISECONDS = 19
    DURATION_MICROSECONDS = 20
    DURATION_NANOSECONDS = 21
    LIST = 24  # results of COLLECT_LIST / COLLECT_SET (ListColumn)


class NullPolicy(enum.IntEnum):
    EXCLUDE = 0
    INCLUDE = 1


class Interpolation(enum.IntEnum):
    """cudf::interpolation (cpp/include/cudf/types.hpp)."""
    LINEAR = 0
    LOWER = 1
    HIGHER = 2
    MIDPOINT = 3
    NEAREST = 4


class NullEquality(enum.IntEnum):
    """cudf::null_equality (cpp/include/cudf/types.hpp)."""
    EQUAL = 0
    UNEQUAL = 1


class NanEquality(enum.IntEnum):
    """cudf::nan_equality (cpp/include/cudf/types.hpp)."""
    ALL_EQUAL = 0
    UNEQUAL = 1


class Sorted(enum.IntEnum):
    NO = 0
    YES = 1


class Kind(enum.IntEnum):
    SUM = 0
    PRODUCT = 1
    MIN = 2
    MAX = 3
    COUNT_VALID = 4
    COUNT_ALL = 5
    ANY = 6
    ALL = 7
    SUM_OF_SQUARES = 8
    MEAN = 9
    M2 = 10
    VARIANCE = 11
    STD = 12
    MEDIAN = 13
    QUANTILE = 14
    ARGMAX = 15
    ARGMIN = 16
    NUNIQUE = 17
    NTH_ELEMENT = 18
    ROW_NUMBER = 19
    RANK = 20
    COLLECT_LIST = 21
    COLLECT_SET = 22


# numpy storage dtype of each fixed-width type
NUMPY_DTYPE = {
    TypeId.INT8: np.int8, TypeId.INT16: np.int16, TypeId.INT32: np.int32, TypeId.INT64: np.int64,
    TypeId.UINT8: np.uint8, TypeId.UINT16: np.uint16, TypeId.UINT32: np.uint32, TypeId.UINT64: np.uint64,
    TypeId.FLOAT32: np.float32, TypeId.FLOAT64: np.float64, TypeId.BOOL8: np.uint8,
    TypeId.TIMESTAMP_DAYS: np.int32, TypeId.TIMESTAMP_SECONDS: np.int64,
    TypeId.TIMESTAMP_MILLISECONDS: np.int64, TypeId.TIMESTAMP_MICROSECONDS: np.int64,
    TypeId.TIMESTAMP_NANOSECONDS: np.int64, TypeId.DURATION_DAYS: np.int32,
    TypeId.DURATION_SECONDS: np.int64, TypeId.DURATION_MILLISECONDS: np.int64,
    TypeId.DURATION_MICROSECONDS: np.int64, TypeId.DURATION_NANOSECONDS: np.int64,
}

FROM_NUMPY = {
    np.dtype(np.int8): TypeId.INT8, np.dtype(np.int16): TypeId.INT16, np.dtype(np.int32): TypeId.INT32,
    np.dtype(np.int64): TypeId.INT64, np.dtype(np.uint8): TypeId.UINT8, np.dtype(np.uint16): TypeId.UINT16,
    np.dtype(np.uint32): TypeId.UINT32, np.dtype(np.uint64): TypeId.UINT64,
    np.dtype(np.float32): TypeId.FLOAT32, np.dtype(np.float64): TypeId.FLOAT64,
    np.dtype(np.bool_): TypeId.BOOL8,
}


def type_size(t: int) -> int:
    return np.dtype(NUMPY_DTYPE[TypeId(t)]).itemsize
