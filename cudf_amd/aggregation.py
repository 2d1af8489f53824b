"""Aggregation factories named as in pylibcudf.aggregation
(python/cudf/cudf/_lib/pylibcudf/aggregation.pyx:143-367)."""
from __future__ import annotations

from dataclasses import dataclass

from .types import Interpolation, Kind, NanEquality, NullEquality, NullPolicy


@dataclass(frozen=True)
class Aggregation:
    kind: Kind
    ddof: int = 0              # VARIANCE / STD: ddof; NTH_ELEMENT: n; COLLECT_SET: null equality
    null_policy: int = NullPolicy.EXCLUDE
    interpolation: int = Interpolation.LINEAR  # QUANTILE; COLLECT_SET: NaN equality
    quantile: float = 0.0

    def __repr__(self) -> str:
        return f"Aggregation({self.kind.name})"

    def to_c(self):
        from . import _lib

        return _lib.mgb_aggregation(int(self.kind), int(self.ddof), int(self.null_policy), int(self.interpolation),
                                    float(self.quantile))


def sum() -> Aggregation:  # noqa: A001 - mirrors the reference name
    return Aggregation(Kind.SUM)


def product() -> Aggregation:
    return Aggregation(Kind.PRODUCT)


def min() -> Aggregation:  # noqa: A001
    return Aggregation(Kind.MIN)


def max() -> Aggregation:  # noqa: A001
    return Aggregation(Kind.MAX)


def count(null_handling: NullPolicy = NullPolicy.EXCLUDE) -> Aggregation:
    """COUNT(EXCLUDE) counts valid values, COUNT(INCLUDE) counts rows."""
    return Aggregation(Kind.COUNT_ALL if null_handling == NullPolicy.INCLUDE else Kind.COUNT_VALID)


def sum_of_squares() -> Aggregation:
    return Aggregation(Kind.SUM_OF_SQUARES)


def mean() -> Aggregation:
    return Aggregation(Kind.MEAN)


def m2() -> Aggregation:
    return Aggregation(Kind.M2)


def variance(ddof: int = 1) -> Aggregation:
    return Aggregation(Kind.VARIANCE, ddof)


def std(ddof: int = 1) -> Aggregation:
    return Aggregation(Kind.STD, ddof)


def median() -> Aggregation:
    return Aggregation(Kind.MEDIAN)


def quantile(q: float, interp: Interpolation = Interpolation.LINEAR) -> Aggregation:
    """One quantile per aggregation (pylibcudf takes a list: one aggregation per entry here)."""
    return Aggregation(Kind.QUANTILE, interpolation=int(interp), quantile=float(q))


def nunique(null_handling: NullPolicy = NullPolicy.EXCLUDE) -> Aggregation:
    return Aggregation(Kind.NUNIQUE, null_policy=int(null_handling))


def nth_element(n: int, null_handling: NullPolicy = NullPolicy.INCLUDE) -> Aggregation:
    return Aggregation(Kind.NTH_ELEMENT, ddof=int(n), null_policy=int(null_handling))


def argmax() -> Aggregation:
    return Aggregation(Kind.ARGMAX)


def argmin() -> Aggregation:
    return Aggregation(Kind.ARGMIN)


def collect_list(null_handling: NullPolicy = NullPolicy.INCLUDE) -> Aggregation:
    """Each group's values in row order, as a list (nulls kept as null elements under INCLUDE)."""
    return Aggregation(Kind.COLLECT_LIST, null_policy=int(null_handling))


def collect_set(null_handling: NullPolicy = NullPolicy.INCLUDE, nulls_equal: NullEquality = NullEquality.EQUAL,
                nans_equal: NanEquality = NanEquality.ALL_EQUAL) -> Aggregation:
    """Each group's distinct values, ascending, then its nulls (one when nulls are equal, each otherwise,
    none under EXCLUDE)."""
    return Aggregation(Kind.COLLECT_SET, ddof=int(nulls_equal), null_policy=int(null_handling),
                       interpolation=int(nans_equal))
