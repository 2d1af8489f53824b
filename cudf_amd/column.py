"""Device columns and tables for the Python binding (the pylibcudf ``Column`` /
``Table`` roles). Device memory is held in PyTorch tensors; PyTorch is only
the allocator and stream provider here."""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np
import torch

from .types import FROM_NUMPY, NUMPY_DTYPE, TypeId, type_size


def bitmask_allocation_words(n: int) -> int:
    """32-bit words of a cudf bitmask for n rows, padded to 64 bytes."""
    nbytes = ((n + 31) // 32) * 4
    nbytes = (nbytes + 63) // 64 * 64
    return nbytes // 4


def pack_validity(valid: np.ndarray) -> np.ndarray:
    """bool[n] -> uint32 words, LSB first, bit = 1 valid (cudf bitmask layout)."""
    valid = np.asarray(valid, dtype=bool)
    n = valid.shape[0]
    words = np.zeros(bitmask_allocation_words(n), dtype=np.uint32)
    if n:
        bits = np.packbits(valid, bitorder="little")
        raw = np.zeros(words.size * 4, dtype=np.uint8)
        raw[: bits.size] = bits
        words = raw.view(np.uint32).copy()
    return words


def unpack_validity(words: np.ndarray, n: int, offset: int = 0) -> np.ndarray:
    raw = np.asarray(words, dtype=np.uint32).view(np.uint8)
    bits = np.unpackbits(raw, bitorder="little").astype(bool)
    return bits[offset: offset + n]


class Column:
    """A fixed-width device column: typed data, optional validity bitmask."""

    def __init__(self, type_id: int, size: int, data: Optional[torch.Tensor],
                 mask: Optional[torch.Tensor] = None, null_count: int = 0, offset: int = 0):
        self.type_id = TypeId(type_id)
        self.size = int(size)
        self.data = data
        self.mask = mask
        self.null_count = int(null_count) if mask is not None else 0
        self.offset = int(offset)
        if data is not None and size:
            need = (offset + size) * type_size(self.type_id)
            if data.numel() * data.element_size() < need:
                raise ValueError("data buffer is smaller than offset + size elements")
        if mask is not None and mask.dtype != torch.int32:
            raise ValueError("mask must be an int32 tensor of bitmask words")

    @property
    def data_ptr(self) -> int:
        return 0 if self.data is None else self.data.data_ptr()

    @property
    def mask_ptr(self) -> int:
        return 0 if self.mask is None else self.mask.data_ptr()

    def nullable(self) -> bool:
        return self.mask is not None

    @staticmethod
    def from_numpy(values: np.ndarray, valid: Optional[np.ndarray] = None, type_id: Optional[int] = None,
                   device: str = "cuda") -> "Column":
        values = np.ascontiguousarray(values)
        if type_id is None:
            type_id = FROM_NUMPY[values.dtype]
        type_id = TypeId(type_id)
        storage = values.astype(NUMPY_DTYPE[type_id], copy=False)
        data = torch.from_numpy(storage.view(np.uint8).copy()).to(device) if storage.size else \
            torch.empty(0, dtype=torch.uint8, device=device)
        mask = None
        null_count = 0
        if valid is not None:
            valid = np.asarray(valid, dtype=bool)
            mask = torch.from_numpy(pack_validity(valid).view(np.int32)).to(device)
            null_count = int((~valid).sum())
        return Column(type_id, storage.size, data, mask, null_count)

    @staticmethod
    def from_torch(t: torch.Tensor, type_id: int, mask: Optional[torch.Tensor] = None,
                   null_count: int = 0) -> "Column":
        if not t.is_contiguous():
            raise ValueError("column data must be contiguous")
        flat = t.reshape(-1)
        n = flat.numel() * flat.element_size() // type_size(type_id)
        return Column(type_id, n, flat, mask, null_count)

    def to_numpy(self):
        """(values, valid-or-None) copied to the host."""
        dt = np.dtype(NUMPY_DTYPE[self.type_id])
        if self.size == 0 or self.data is None:
            vals = np.zeros(0, dtype=dt)
        else:
            raw = self.data.reshape(-1).view(torch.uint8).cpu().numpy()
            start = self.offset * dt.itemsize
            vals = raw[start: start + self.size * dt.itemsize].view(dt).copy()
        valid = None
        if self.mask is not None:
            valid = unpack_validity(self.mask.cpu().numpy().view(np.uint32), self.size, self.offset)
        return vals, valid

    def __repr__(self) -> str:
        return f"Column({self.type_id.name}, size={self.size}, null_count={self.null_count})"


class ListColumn:
    """A LIST column (the result of COLLECT_LIST / COLLECT_SET): `size` lists,
    list i holding elements offsets[i]..offsets[i+1] of `child`; lists are
    never null themselves (cudf's lists column)."""

    type_id = TypeId.LIST

    def __init__(self, size: int, offsets: Column, child: Column):
        self.size = int(size)
        self.offsets = offsets
        self.child = child
        self.mask = None
        self.null_count = 0

    def nullable(self) -> bool:
        return False

    def to_numpy(self):
        """(offsets int32[size + 1], (child values, child valid-or-None)) copied to the host."""
        off, _ = self.offsets.to_numpy()
        if off.size == 0:
            off = np.zeros(1, dtype=np.int32)
        return off, self.child.to_numpy()

    def to_lists(self):
        """Python lists of every row's elements (None for a null element)."""
        off, (vals, valid) = self.to_numpy()
        out = []
        for i in range(self.size):
            a, b = int(off[i]), int(off[i + 1])
            out.append([None if (valid is not None and not valid[j]) else vals[j].item() for j in range(a, b)])
        return out

    def __repr__(self) -> str:
        return f"ListColumn({self.child.type_id.name}, size={self.size}, elements={self.child.size})"


class Table:
    """An ordered set of equal-length columns."""

    def __init__(self, columns: Sequence[Column]):
        self.columns = list(columns)
        sizes = {c.size for c in self.columns}
        if len(sizes) > 1:
            raise ValueError("Column size mismatch.")

    def num_columns(self) -> int:
        return len(self.columns)

    def num_rows(self) -> int:
        return self.columns[0].size if self.columns else 0
