"""GPU closures (gpu/transitive_closure.py:26-652 of the reference).

transitive_closure_dag_gpu: on-device levels, then one launch per level that
closes every vertex of the level (librbg_hip.so: rbg_dag_levels,
rbg_dag_closure_levels_*, rbg_dag_gather_*).  Diameter: floor(log2) of the
largest |final value| above 1 (gpu/transitive_closure.py:639-646).

transitive_closure_gpu: repeated squaring R <- R (+) R^2 with the reference's
smallest-|value| merge, until R stops changing (gpu/transitive_closure.py:
26-74).  Diameter: max(#changing iterations, floor(log2 max |final| > 1)).
"""
from __future__ import annotations

from typing import Tuple

from .. import sparse_dev as sd
from .csr_gpu import CSRMatrixGPU

__all__ = ["transitive_closure_gpu", "transitive_closure_dag_gpu", "sparse_avos_sum_gpu", "sparse_equal_gpu"]


def sparse_avos_sum_gpu(A: CSRMatrixGPU, B: CSRMatrixGPU) -> CSRMatrixGPU:
    if A.shape != B.shape:
        raise ValueError(f"Shape mismatch: {A.shape} vs {B.shape}")
    if A.nnz == 0 and B.nnz == 0:
        return A.copy()
    if A.nnz == 0:
        return B.copy()
    if B.nnz == 0:
        return A.copy()
    C = sd.avos_sum_csr(sd.canonicalize(A._dev()), sd.canonicalize(B._dev()), min_abs=True)
    return CSRMatrixGPU._from_dev(C, triangular=A.triangular and B.triangular)


def sparse_equal_gpu(A: CSRMatrixGPU, B: CSRMatrixGPU) -> bool:
    if A.shape != B.shape or A.nnz != B.nnz:
        return False
    if A.nnz == 0:
        return True
    return sd.csr_equal(A._dev(), B._dev())


def _final_diameter(C: sd.DevCSR) -> int:
    """floor(log2 max |value|) when that maximum exceeds 1, else 0
    (gpu/transitive_closure.py:43-50 of the reference): rbg_csr_max_abs."""
    from ..dense import status
    from .. import _lib
    if C.nnz == 0:
        return 0
    st = status()
    st.reset()
    sfx = "i32" if C.data.dtype == _lib.torch().int32 else "i64"
    _lib.call(f"rbg_csr_max_abs_{sfx}", _lib.ptr(C.indptr), _lib.ptr(C.data), C.n_rows, st.p, _lib.stream_handle())
    m = int(st.read().aux[0])
    return m.bit_length() - 1 if m > 1 else 0


def transitive_closure_gpu(A: CSRMatrixGPU, max_iterations: int = 64) -> Tuple[CSRMatrixGPU, int]:
    if A.nnz == 0:
        return A.copy(), 0
    R = sd.canonicalize(A._dev())
    tri = A.triangular
    diameter = 0
    for iteration in range(max_iterations):
        R2, _ = sd.spgemm(R, R, mask_upper=tri)
        Rn = sd.avos_sum_csr(R, R2, min_abs=True)
        if sd.csr_equal(R, Rn):
            break
        R = Rn
        diameter = iteration + 1
    d = _final_diameter(R)
    if d:
        diameter = max(diameter, d)
    return CSRMatrixGPU._from_dev(R, triangular=tri), diameter


def transitive_closure_dag_gpu(A: CSRMatrixGPU) -> Tuple[CSRMatrixGPU, int]:
    from ..sparse.csgraph.cycleerror import CycleError
    if A.nnz == 0:
        return A.copy(), 0
    C, info = sd.dag_closure(sd.canonicalize(A._dev()))
    if C is None:
        raise CycleError("Graph contains a cycle", vertex=info.cycle_vertex)
    if info.events:
        raise OverflowError("AVOS product overflow in the DAG closure")
    return CSRMatrixGPU._from_dev(C, triangular=A.triangular), info.diameter_final
