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
        from .transitive_closure import transitive_closure_dijkstra
        return transitive_closure_dijkstra(R)
    raise ValueError(f"unknown method {method!r}")
