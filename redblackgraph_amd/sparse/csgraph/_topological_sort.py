"""Topological sort of a DAG and the triangularity test
(sparse/csgraph/_topological_sort.pyx:30-265 of the reference).

The order comes from the device's Kahn levels (rbg_topological_sort): every
vertex is placed after the vertices of its longest path to a sink, and the
sort lists the vertices by decreasing level, so each edge i -> j (a nonzero
A[i, j], i != j) has i before j, the property the reference's depth-first
order has.  Both are topological orders of the graph; where a level holds
several vertices their order differs from the depth-first one.
"""
from __future__ import annotations

from typing import Dict

import numpy as np
import scipy.sparse as sp

from ... import _lib
from ...types import Ordering
from .cycleerror import CycleError


def _as_csr(A) -> sp.csr_matrix:
    """The reference's conversion: dense input through csr_matrix(A, int32),
    other sparse formats through tocsr() (_topological_sort.pyx:77-82)."""
    if not sp.isspmatrix(A):
        return sp.csr_matrix(np.asarray(A), dtype=np.int32)
    return A if sp.isspmatrix_csr(A) else A.tocsr()


def _device_structure(A: sp.csr_matrix):
    t = _lib.torch()
    dev = _lib.device()
    indptr = t.from_numpy(np.ascontiguousarray(A.indptr, dtype=np.int64)).to(dev)
    indices = t.from_numpy(np.ascontiguousarray(A.indices, dtype=np.int32)).to(dev)
    return indptr, indices


def topological_levels(A):
    """(order, level): the vertices by decreasing Kahn level, and each
    vertex's level (the longest path from it to a vertex without non-self
    successors).  Raises CycleError("Graph contains a cycle", vertex=...)."""
    from ...dense import status
    A = _as_csr(A)
    n = A.shape[0]
    if n == 0:
        return np.array([], dtype=np.int32), np.array([], dtype=np.int32)
    t = _lib.torch()
    indptr, indices = _device_structure(A)
    ws = t.empty(int(_lib.load().rbg_dag_workspace_bytes(n, int(A.nnz))), dtype=t.uint8, device=indptr.device)
    order = t.empty(n, dtype=t.int32, device=indptr.device)
    level = t.empty(n, dtype=t.int32, device=indptr.device)
    st = status()
    st.reset()
    _lib.call("rbg_topological_sort", _lib.ptr(indptr), _lib.ptr(indices), n, _lib.ptr(ws), ws.numel(),
              _lib.ptr(level), _lib.ptr(order), st.p, _lib.stream_handle())
    s = st.read()
    if s.flags & _lib.F_CYCLE:
        raise CycleError("Graph contains a cycle", vertex=int(s.aux[0]))
    if s.flags & _lib.F_CAPACITY or not s.aux[7]:
        raise RuntimeError("rbg_topological_sort did not place every vertex")
    return order.cpu().numpy(), level.cpu().numpy()


def topological_sort(A) -> np.ndarray:
    """_topological_sort.pyx:30-163: permutation p (p[new] = old, int32) such
    that every edge i -> j of A satisfies new(i) < new(j); self-loops are
    ignored.  Raises CycleError if the graph has a cycle."""
    return topological_levels(A)[0]


def topological_ordering(A) -> Ordering:
    """_topological_sort.pyx:166-219: the sort applied to A (upper triangular
    result, the triangular filter on), with the component sizes."""
    from ._components import find_components, find_components_sparse
    from ._ordering import permute, permute_sparse
    perm = topological_sort(A)
    q: Dict[int, int] = {}
    if sp.isspmatrix(A):
        find_components_sparse(A, q)
        permuted = permute_sparse(A, perm, assume_upper_triangular=True)
    else:
        find_components(A, q)
        permuted = permute(A, perm, assume_upper_triangular=True)
    return Ordering(permuted, perm, q)


def is_upper_triangular(A, strict: bool = False) -> bool:
    """_topological_sort.pyx:222-265: no stored entry below the diagonal
    (strict: none on it either).  Counts run on the device over the CSR
    structure, as the reference walks it (stored zeros count)."""
    from ...dense import status
    A = _as_csr(A)
    n = A.shape[0]
    if n == 0:
        return True
    indptr, indices = _device_structure(A)
    st = status()
    st.reset()
    _lib.call("rbg_csr_triangles", _lib.ptr(indptr), _lib.ptr(indices), n, st.p, _lib.stream_handle())
    s = st.read()
    below, diag = int(s.aux[0]), int(s.aux[1])
    return below == 0 and (not strict or diag == 0)
