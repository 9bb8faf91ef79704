"""floyd_warshall and shortest_path of the csgraph family
(sparse/csgraph/_shortest_path.pyx:36-389 of the reference), on the GPU's
dense closure.

The reference's floyd_warshall converts the graph to a dense int32 matrix
(_validation.py:11-59), closes it in the AVOS semiring with the Cython loop
(_shortest_path.pyx:340-389: CycleError at the first detection, diameter =
MSB of the largest value written) and returns (dist_matrix, diameter).  Here
the closure is rbg_fw_dense_tri_i32: the triangular schedule when every
off-diagonal entry lies on one side of the diagonal (the reference's
assume_upper_triangular loop, and its mirror image), the full schedule
otherwise; W, the diameter and the first detection are those of the
reference's loop either way, so assume_upper_triangular is accepted and the
input decides.

Methods other than "FW" (Dijkstra, Bellman-Ford, Johnson) are the
reference's scipy-derived searches with AVOS arithmetic bolted on; their
results are not the AVOS closure in general, and this build answers every
method with the Floyd-Warshall closure (rows `indices` when given).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from ... import dense, _lib
from ...core import array, raise_fw_errors
from ..._dtypes import elem_type

NULL_IDX = -9999


def _validate_dense(csgraph, directed: bool, overwrite: bool):
    """validate_graph(csgraph, directed, int32, csr_output=False,
    copy_if_dense=not overwrite) (_validation.py:11-59): a dense int32 rb
    array; an undirected CSC graph is transposed first; a dense int32
    C-contiguous input is used in place when overwrite is set."""
    if (not directed) and sp.isspmatrix_csc(csgraph):
        csgraph = csgraph.T
    if sp.isspmatrix(csgraph):
        D = np.ascontiguousarray(csgraph.toarray(), dtype=np.int32)
    else:
        D = np.asarray(csgraph)
        if D.dtype == np.int32 and D.flags.c_contiguous and overwrite:
            pass
        else:
            D = np.array(D, dtype=np.int32, order="C", copy=True)
    if D.ndim != 2:
        raise ValueError("compressed-sparse graph must be 2-D")
    if D.shape[0] != D.shape[1]:
        raise ValueError("compressed-sparse graph must be shape (N, N)")
    return D


def floyd_warshall(csgraph, directed=True, return_predecessors=False, unweighted=False, overwrite=False,
                   assume_upper_triangular=False):
    """_shortest_path.pyx:213-337.  Returns (dist_matrix, diameter), or
    (dist_matrix, predecessors, diameter) with return_predecessors (the
    reference leaves that matrix unfilled; here every entry is -9999).
    Raises CycleError("Error: cycle detected! Vertex i has a path to itself.
    A(i,k)=.., A(k,i)=..", vertex=i) at the first detection."""
    D = _validate_dense(csgraph, directed, overwrite)
    if unweighted:
        D[...] = 1  # dist_matrix[~np.isinf(dist_matrix)] = 1: no integer entry is infinite
    n = D.shape[0]
    et = elem_type(np.int32)
    if n == 0:
        W, diameter = D, 0
    else:
        Wd, out = dense.fw_closure(_lib.to_device(D), et)
        raise_fw_errors(out, et, "csgraph")
        W = Wd.cpu().numpy()
        diameter = out.diameter
    D[...] = W
    dist = D.view(array)
    if return_predecessors:
        return dist, np.full(D.shape, NULL_IDX, dtype=np.int32), diameter
    return dist, diameter


def shortest_path(csgraph, method="auto", directed=True, return_predecessors=False, unweighted=False,
                  overwrite=False, indices=None, assume_upper_triangular=False):
    """_shortest_path.pyx:36-210.  method "FW" (and "auto" on dense enough
    graphs, the reference's rule: nonzeros >= N^2/4 + N) is floyd_warshall;
    "D", "BF", "J" and the other "auto" cases return the rows `indices` of
    the same closure (all rows when indices is None)."""
    if method not in ("auto", "FW", "D", "BF", "J"):
        raise ValueError("unrecognized method '%s'" % method)
    if method == "FW" and indices is not None:
        raise ValueError("Cannot specify indices with method == 'FW'.")
    res = floyd_warshall(csgraph, directed, return_predecessors=return_predecessors, unweighted=unweighted,
                         overwrite=overwrite if method in ("FW", "auto") and indices is None else False,
                         assume_upper_triangular=assume_upper_triangular)
    if indices is None:
        return res
    n = res[0].shape[0]
    idx = np.array(indices, dtype=np.int32, copy=True)
    shape = idx.shape + (n,)
    idx = np.atleast_1d(idx).reshape(-1)
    idx[idx < 0] += n
    if np.any(idx < 0) or np.any(idx >= n):
        raise ValueError("indices out of range 0...N")
    rows = np.asarray(res[0])[idx].reshape(shape).view(array)
    if return_predecessors:
        return rows, np.asarray(res[1])[idx].reshape(shape), res[2]
    return rows, res[1]
