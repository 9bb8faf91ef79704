"""CSRMatrixGPU: a CSR matrix resident in MI355X HBM
(gpu/csr_gpu.py:32-325 of the reference).

Attributes mirror the reference (``data``, ``indices``, ``indptr``, ``shape``,
``triangular``).  Storage: torch tensors on the device; ``indptr`` is int64,
``indices`` int32, ``data`` int32 by default (the reference casts to int32) or
int64 when int64 values are given, which the 1M-vertex workload uses.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from .. import _lib
from .. import sparse_dev as sd


def _as_dev(x, dtype):
    t = _lib.torch()
    if isinstance(x, t.Tensor):
        return x.to(device=_lib.device(), dtype=dtype).contiguous()
    return t.from_numpy(np.ascontiguousarray(np.asarray(x))).to(device=_lib.device(), dtype=dtype)


class CSRMatrixGPU:
    def __init__(self, data, indices, indptr, shape: Tuple[int, int], *, triangular: bool = False,
                 validate: bool = True):
        t = _lib.torch()
        if isinstance(data, t.Tensor):
            ddt = t.int64 if data.dtype == t.int64 else t.int32
        else:
            ddt = t.int64 if np.asarray(data).dtype == np.int64 else t.int32
        self.data = _as_dev(data, ddt)
        self.indices = _as_dev(indices, t.int32)
        self.indptr = _as_dev(indptr, t.int64)
        self.shape = tuple(int(s) for s in shape)
        self.triangular = triangular
        if validate:
            self._validate_csr()
            if triangular:
                self._validate_triangular()

    # -- construction / transfer -------------------------------------------
    @classmethod
    def _from_dev(cls, c: sd.DevCSR, triangular=False) -> "CSRMatrixGPU":
        m = cls.__new__(cls)
        m.data, m.indices, m.indptr, m.shape, m.triangular = c.data, c.indices, c.indptr, tuple(c.shape), triangular
        return m

    def _dev(self) -> sd.DevCSR:
        return sd.DevCSR(self.indptr, self.indices, self.data, self.shape)

    @classmethod
    def from_cpu(cls, cpu_matrix, *, triangular: Optional[bool] = None, validate: bool = True) -> "CSRMatrixGPU":
        import scipy.sparse as sp
        if not sp.isspmatrix_csr(cpu_matrix):
            cpu_matrix = sp.csr_matrix(cpu_matrix)
        dt = np.int64 if cpu_matrix.dtype == np.int64 else np.int32
        m = cls(cpu_matrix.data.astype(dt), cpu_matrix.indices, cpu_matrix.indptr, cpu_matrix.shape,
                triangular=False, validate=validate)
        if triangular is None:
            triangular = m._is_upper_triangular()
        elif triangular and validate:
            m._validate_triangular()
        m.triangular = bool(triangular)
        return m

    def to_cpu(self):
        return sd.to_scipy(self._dev())

    # -- checks (torch reductions on the device) ------------------------------
    def _row_ids(self):
        t = _lib.torch()
        return t.repeat_interleave(t.arange(self.shape[0], device=self.indices.device),
                                   self.indptr[1:] - self.indptr[:-1])

    def _validate_csr(self):
        t = _lib.torch()
        rows, cols = self.shape
        nnz = self.data.numel()
        if self.indices.numel() != nnz:
            raise ValueError(f"indices length {self.indices.numel()} != data length {nnz}")
        if self.indptr.numel() != rows + 1:
            raise ValueError(f"indptr length {self.indptr.numel()} != rows + 1 ({rows + 1})")
        if rows and bool((self.indptr[1:] < self.indptr[:-1]).any()):
            raise ValueError("indptr is not monotonically non-decreasing")
        first, last = int(self.indptr[0].item()), int(self.indptr[-1].item())
        if first != 0:
            raise ValueError(f"indptr[0] should be 0, got {first}")
        if last != nnz:
            raise ValueError(f"indptr[-1] should be {nnz}, got {last}")
        if nnz and bool(((self.indices < 0) | (self.indices >= cols)).any()):
            raise ValueError(f"Column indices out of bounds [0, {cols})")

    def _is_upper_triangular(self) -> bool:
        if self.nnz == 0:
            return True
        return not bool((self.indices < self._row_ids()).any())

    def _validate_triangular(self):
        if self.nnz == 0:
            return
        t = _lib.torch()
        r = self._row_ids()
        bad = self.indices < r
        if bool(bad.any()):
            k = int(t.nonzero(bad)[0, 0].item())
            row = int(r[k].item())
            a, b = int(self.indptr[row].item()), int(self.indptr[row + 1].item())
            min_col = int(self.indices[a:b].min().item())
            raise ValueError(f"Not upper triangular: row {row} has column {min_col} < {row}")
        same = r[1:] == r[:-1]
        desc = same & (self.indices[1:] < self.indices[:-1])
        if bool(desc.any()):
            k = int(t.nonzero(desc)[0, 0].item())
            raise ValueError(f"Columns not sorted in row {int(r[k].item())}")

    # -- properties -----------------------------------------------------------
    @property
    def nnz(self) -> int:
        return int(self.data.numel())

    @property
    def dtype(self):
        return self.data.dtype

    def __matmul__(self, other):
        from .spgemm import spgemm
        if not isinstance(other, CSRMatrixGPU):
            return NotImplemented
        return spgemm(self, other)

    def copy(self) -> "CSRMatrixGPU":
        m = CSRMatrixGPU.__new__(CSRMatrixGPU)
        m.data, m.indices, m.indptr = self.data.clone(), self.indices.clone(), self.indptr.clone()
        m.shape, m.triangular = self.shape, self.triangular
        return m

    def eliminate_zeros(self):
        """Remove explicit zeros in place (device canonicalisation)."""
        if self.nnz == 0 or not bool((self.data == 0).any()):
            return
        c = sd.canonicalize(self._dev())
        self.data, self.indices, self.indptr = c.data, c.indices, c.indptr

    def transitive_closure(self, max_iterations=64):
        from .transitive_closure import transitive_closure_gpu
        return transitive_closure_gpu(self, max_iterations=max_iterations)

    def prefetch(self, device=None):
        """No-op: the arrays already live in device HBM (the reference prefetches
        unified memory, csr_gpu.py:276-297)."""
        return None

    def __repr__(self) -> str:
        tri = " (upper triangular)" if self.triangular else ""
        return f"CSRMatrixGPU({self.shape[0]}x{self.shape[1]}, nnz={self.nnz}{tri})"

    def memory_usage(self) -> dict:
        d = self.data.numel() * self.data.element_size()
        i = self.indices.numel() * self.indices.element_size()
        p = self.indptr.numel() * self.indptr.element_size()
        tot = d + i + p
        return {"data": d, "indices": i, "indptr": p, "total": tot, "total_mb": tot / 1024 ** 2,
                "total_gb": tot / 1024 ** 3}


def validate_triangular_mask(i: int, j: int) -> bool:
    return j >= i
