"""Aggregation factories named as in pylibcudf.aggregation
(python/cudf/cudf/_lib/pylibcudf/aggregation.pyx:143-367)."""
from __future__ import annotations

from dataclasses import dataclass

from .types import Kind, NullPolicy


@dataclass(frozen=True)
class Aggregation:
    kind: Kind
    ddof: int = 0

    def __repr__(self) -> str:
        return f"Aggregation({self.kind.name})"


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


def argmax() -> Aggregation:
    return Aggregation(Kind.ARGMAX)


def argmin() -> Aggregation:
    return Aggregation(Kind.ARGMIN)
