"""csgraph closure entry points (sparse/csgraph/transitive_closure.py:17-574)."""
from __future__ import annotations

import numpy as np

from ... import dense, _lib
from ...core import array, raise_fw_errors
from ..._dtypes import elem_type
from ...types import TransitiveClosure


def _as_dense_int32(R) -> np.ndarray:
    """validate_graph(csgraph, directed, DTYPE, csr_output=False): dense int32
    (sparse/csgraph/_validation.py:8,38-46)."""
    if hasattr(R, "toarray"):
        R = R.toarray()
    return np.ascontiguousarray(np.asarray(R).astype(np.int32))


def transitive_closure_floyd_warshall(R, assume_upper_triangular=False) -> TransitiveClosure:
    """Floyd-Warshall closure with the Cython path's rules
    (_shortest_path.pyx:340-389): int32, CycleError at the first detection.
    ``assume_upper_triangular`` selects the i<=k<=j loop in the reference; the
    values and diameter it produces equal the full schedule's on the inputs it
    is defined for, so the GPU runs the full schedule either way."""
    A = _as_dense_int32(R)
    et = elem_type(np.int32)
    if A.shape[0] == 0:
        return TransitiveClosure(array(A), 0)
    W, out = dense.fw_closure(_lib.to_device(A), et)
    raise_fw_errors(out, et, "csgraph")
    return TransitiveClosure(array(W.cpu().numpy()), out.diameter)


def transitive_closure(R, method="D", assume_upper_triangular=False) -> TransitiveClosure:
    if method == "FW":
        return transitive_closure_floyd_warshall(R, assume_upper_triangular)
    if method == "D":
        return transitive_closure_dijkstra(R)
    raise ValueError(f"unknown method {method!r}")


class DensificationError(Exception):
    """_density.py of the reference: raised when sparse-only mode meets a
    graph that would need O(N^2) memory (a cycle)."""

    def __init__(self, message, density=None, threshold=None):
        super().__init__(message)
        self.density = density
        self.threshold = threshold


def _csr(A):
    from scipy.sparse import csr_matrix, isspmatrix
    if not isspmatrix(A):
        return csr_matrix(np.asarray(A))
    return A.tocsr()


def _is_upper(A) -> bool:
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    return bool(np.all(A.indices >= rows)) if A.nnz else True


def _is_lower(A) -> bool:
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    return bool(np.all(A.indices <= rows)) if A.nnz else True


def transitive_closure_dag_sparse(A) -> TransitiveClosure:
    """Sparse DAG closure (_transitive_closure_dag.pyx:168-351) on the GPU.

    Values are computed in the input's width (int32 for inputs of 32 bits or
    fewer, as the Cython DTYPE; int64 otherwise).  The diameter is the Cython
    rule: MSB of the largest |candidate| (direct edge or nonzero product).
    Raises CycleError("transitive_closure_dag_sparse requires a DAG (no
    cycles)", vertex=...) on a cycle."""
    from ..rbm import rb_matrix
    from ... import sparse_dev as sd
    from .cycleerror import CycleError
    A = _csr(A)
    n = A.shape[0]
    if n == 0:
        from scipy.sparse import csr_matrix
        return TransitiveClosure(csr_matrix((0, 0), dtype=np.int32), 0)
    dt = np.int64 if A.dtype.itemsize > 4 else np.int32
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, dt)))
    if C is None:
        raise CycleError("transitive_closure_dag_sparse requires a DAG (no cycles)", vertex=info.cycle_vertex)
    if info.events:
        raise OverflowError("Avos product results in an overflow.")
    c = sd.to_scipy(C)
    return TransitiveClosure(rb_matrix((c.data, c.indices, c.indptr), shape=c.shape), info.diameter_candidates)


def transitive_closure_squaring(A, max_iterations=64) -> TransitiveClosure:
    """sparse/csgraph/transitive_closure.py:198-269: triangular inputs go to the
    DAG closure; others are squared R <- R (+) R^2 (smallest-|value| merge)
    until stable.  Diameter: floor(log2(max |final|)) when above 1."""
    from ..rbm import rb_matrix
    from ... import sparse_dev as sd
    A = _csr(A)
    n = A.shape[0]
    if n == 0:
        from scipy.sparse import csr_matrix
        return TransitiveClosure(csr_matrix((0, 0), dtype=np.int32), 0)
    if _is_upper(A) or _is_lower(A):
        return transitive_closure_dag_sparse(A)
    R = sd.canonicalize(sd.from_scipy(A, np.int32))
    for _ in range(max_iterations):
        R2, _ev = sd.spgemm(R, R)
        Rn = sd.avos_sum_csr(R, R2, min_abs=True)
        if sd.csr_equal(R, Rn):
            break
        R = Rn
    c = sd.to_scipy(R)
    diameter = 0
    if c.nnz:
        mx = int(np.abs(c.data).max())
        diameter = mx.bit_length() - 1 if mx > 1 else 0
    return TransitiveClosure(rb_matrix((c.data, c.indices, c.indptr), shape=c.shape), diameter)


def transitive_closure_dijkstra(R) -> TransitiveClosure:
    """transitive_closure.py:62-82: the reference routes upper-triangular input
    to the triangular Floyd-Warshall and everything else to a Dijkstra that is
    not AVOS-exact; here both go to the (exact) Floyd-Warshall closure."""
    return transitive_closure_floyd_warshall(R, assume_upper_triangular=True)


def transitive_closure_adaptive(A, method: str = "auto", density_threshold: float = 0.1,
                                component_threshold: int = 1, size_threshold: int = 1000,
                                sparse_only: bool = False) -> TransitiveClosure:
    """transitive_closure.py:441-574 (method selection)."""
    from .cycleerror import CycleError
    A = _csr(A)
    n = A.shape[0]
    if n == 0:
        from scipy.sparse import csr_matrix
        return TransitiveClosure(csr_matrix((0, 0), dtype=np.int32), 0)
    if sparse_only:
        try:
            return transitive_closure_dag_sparse(A)
        except CycleError as e:
            raise DensificationError(f"sparse_only=True requires a DAG, but graph contains cycles: {e}",
                                     density=1.0, threshold=0.0)
    if method == "FW":
        return transitive_closure(A, method="FW")
    if method == "D":
        return transitive_closure(A, method="D")
    from ._components import component_wise_closure, find_components_sparse
    if method == "component":
        return component_wise_closure(A)
    if method == "squaring":
        return transitive_closure_squaring(A)
    if method == "dag_sparse":
        return transitive_closure_dag_sparse(A)
    q = {}
    find_components_sparse(A, q)
    if len(q) > component_threshold:
        return component_wise_closure(A)
    if A.nnz / (n * n) < density_threshold and n > size_threshold:
        return transitive_closure(A, method="D")
    if _is_upper(A) or _is_lower(A):
        return transitive_closure_dag_sparse(A)
    return transitive_closure(A, method="FW")
