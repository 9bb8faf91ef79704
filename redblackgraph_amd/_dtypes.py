"""Element types and the widened device domain.

The reference's C core handles ten integer element types (rbg_math.c.in:15-26);
RED_ONE is -1 for signed types and the all-ones maximum for unsigned ones
(constants.py:283-313).  On the device every type is held in a signed
"widened" domain -- int32 for widths 8/16 and signed 32, int64 for unsigned 32
and 64 -- with RED_ONE = -1, and the kernels receive the element width and
signedness so they apply the element type's overflow rules.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class ElemType:
    dtype: np.dtype
    width: int
    unsigned: bool
    dev: np.dtype  # widened device dtype

    @property
    def native(self) -> bool:
        """True when the fast kernels apply (signed, device width == element width)."""
        return (not self.unsigned) and self.width == self.dev.itemsize * 8

    @property
    def red(self) -> int:
        return (1 << self.width) - 1 if self.unsigned else -1


def elem_type(dtype) -> ElemType:
    dt = np.dtype(dtype)
    if dt.kind not in "iu":
        raise TypeError(f"AVOS arithmetic needs an integer dtype, got {dt}")
    w = dt.itemsize * 8
    uns = dt.kind == "u"
    if w <= 16 or (w == 32 and not uns):
        dev = np.dtype(np.int32)
    else:
        dev = np.dtype(np.int64)
    return ElemType(dt, w, uns, dev)


def widen(a: np.ndarray, et: ElemType) -> np.ndarray:
    a = np.asarray(a)
    if et.unsigned:
        out = a.astype(et.dev) if et.width < 64 else a.view(np.int64).copy()
        if et.width < 64:
            out[a == np.iinfo(et.dtype).max] = -1
        return out
    return a.astype(et.dev)


def narrow(b: np.ndarray, et: ElemType) -> np.ndarray:
    b = np.asarray(b)
    if et.unsigned:
        if et.width == 64:
            return b.astype(np.int64).view(np.uint64)
        out = b.astype(np.int64)
        out[out == -1] = (1 << et.width) - 1
        return out.astype(et.dtype)
    return b.astype(et.dtype)


def as_unsigned_text(v: int, width: int) -> int:
    """How the reference's %u/%lu formats print an element value."""
    return int(v) & ((1 << width) - 1)


def msb(x: int) -> int:
    return int(x).bit_length() - 1 if x > 1 else 0
