"""AVOS SpGEMM on the GPU (gpu/spgemm.py:32-270 of the reference): one
wavefront per output row folds the scaled rows of B with a merge-path merge
(librbg_hip.so: rbg_spgemm_rows_*).  Zero sums are dropped."""
from __future__ import annotations

import time
from typing import Optional

from .. import sparse_dev as sd
from .csr_gpu import CSRMatrixGPU

__all__ = ["spgemm", "spgemm_upper_triangular", "matmul_gpu", "SpGEMMStats", "spgemm_with_stats"]


def _canonical(A: CSRMatrixGPU) -> sd.DevCSR:
    return sd.canonicalize(A._dev())


def spgemm(A: CSRMatrixGPU, B: CSRMatrixGPU = None, *, upper_triangular: bool = None) -> CSRMatrixGPU:
    self_multiply = B is None or B is A
    if B is None:
        B = A
    if A.shape[1] != B.shape[0]:
        raise ValueError(f"Incompatible shapes: A is {A.shape}, B is {B.shape}. A.shape[1] must equal B.shape[0].")
    if upper_triangular is None:
        upper_triangular = self_multiply and A.triangular
    a = _canonical(A)
    b = a if self_multiply else _canonical(B)
    C, events = sd.spgemm(a, b, mask_upper=bool(upper_triangular))
    return CSRMatrixGPU._from_dev(C, triangular=bool(upper_triangular))


def spgemm_upper_triangular(A: CSRMatrixGPU, *, validate: bool = True) -> CSRMatrixGPU:
    if A.shape[0] != A.shape[1]:
        raise ValueError(f"Matrix must be square, got {A.shape}")
    if validate and not A.triangular:
        raise ValueError("Matrix must be upper triangular")
    return spgemm(A, upper_triangular=True)


def matmul_gpu(A: CSRMatrixGPU, B: Optional[CSRMatrixGPU] = None) -> CSRMatrixGPU:
    return spgemm(A, B)


class SpGEMMStats:
    """Statistics of one SpGEMM (gpu/spgemm.py:172-230)."""

    def __init__(self, input_nnz, output_nnz, input_shape, symbolic_time=0.0, numeric_time=0.0, total_time=0.0):
        self.input_nnz = input_nnz
        self.output_nnz = output_nnz
        self.input_shape = input_shape
        self.symbolic_time = symbolic_time
        self.numeric_time = numeric_time
        self.total_time = total_time

    @property
    def fill_ratio(self) -> float:
        return self.output_nnz / self.input_nnz if self.input_nnz else 0.0

    def __repr__(self):
        return (f"SpGEMMStats(input_nnz={self.input_nnz}, output_nnz={self.output_nnz}, "
                f"fill_ratio={self.fill_ratio:.2f}, total_time={self.total_time * 1e3:.2f}ms)")


def spgemm_with_stats(A: CSRMatrixGPU, B: CSRMatrixGPU = None, *, upper_triangular: bool = None):
    from .. import _lib
    t = _lib.torch()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    C = spgemm(A, B, upper_triangular=upper_triangular)
    t.cuda.synchronize()
    dt = time.perf_counter() - t0
    # one fused pass: symbolic sizing and numeric values happen in the same kernel
    return C, SpGEMMStats(A.nnz, C.nnz, A.shape, 0.0, dt, dt)
