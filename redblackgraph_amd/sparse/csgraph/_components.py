"""Connected components and the per-component closure
(sparse/csgraph/_components.pyx and transitive_closure.py:85-195 of the
reference).

Component finding runs on the GPU (rbg_components: hooking and pointer
jumping).  component_wise_closure extracts every component into one device
pool, closes them all together with the batched Floyd-Warshall launcher, and
merges the closed components back into one CSR matrix on the device.
extract_submatrix and merge_component_matrices keep the reference's host
(scipy) signatures for callers that use them directly.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np
import scipy.sparse as sp

from ... import _lib


def _csr(A) -> sp.csr_matrix:
    if sp.isspmatrix_csr(A):
        return A
    return A.tocsr() if sp.isspmatrix(A) else sp.csr_matrix(np.asarray(A))


def _device_labels(A: sp.csr_matrix):
    """Smallest vertex of each vertex's component (device int32 tensor)."""
    t = _lib.torch()
    dev = _lib.device()
    n = A.shape[0]
    indptr = t.from_numpy(np.ascontiguousarray(A.indptr, dtype=np.int64)).to(dev)
    indices = t.from_numpy(np.ascontiguousarray(A.indices, dtype=np.int32)).to(dev)
    labels = t.empty(max(n, 1), dtype=t.int32, device=dev)
    flags = t.zeros(2, dtype=t.int32, device=dev)
    rounds = ctypes.c_int32(0)
    _lib.call("rbg_components", _lib.ptr(indptr), _lib.ptr(indices), n, _lib.ptr(labels), _lib.ptr(flags),
              ctypes.byref(rounds), _lib.stream_handle())
    return labels[:n]


def _ids_from_labels(roots: np.ndarray, q: Optional[Dict[int, int]]) -> np.ndarray:
    uniq, ids, counts = np.unique(roots, return_inverse=True, return_counts=True)
    if q is not None:
        for cid, c in enumerate(counts):
            q[int(cid)] = int(c)
    return ids.astype(np.uint32)


def find_components_sparse(A, q: Optional[Dict[int, int]] = None) -> np.ndarray:
    """_components.pyx:86-177: component id per vertex (uint32), ids in order
    of each component's smallest vertex; q receives id -> vertex count."""
    A = _csr(A)
    if A.shape[0] == 0:
        return np.zeros(0, dtype=np.uint32)
    return _ids_from_labels(_device_labels(A).cpu().numpy(), q)


def find_components(A, q: Optional[Dict[int, int]] = None) -> np.ndarray:
    """_components.pyx:15-81: the same numbering for dense or sparse input
    (a nonzero in row or column j joins j to the component)."""
    return find_components_sparse(A, q)


def get_component_vertices(component_labels: np.ndarray) -> Dict[int, np.ndarray]:
    """_components.pyx:326-344: component id -> sorted vertex indices (int32)."""
    labels = np.asarray(component_labels)
    order = np.argsort(labels, kind="stable")
    uniq, starts = np.unique(labels[order], return_index=True)
    bounds = list(starts) + [labels.size]
    return {int(u): order[bounds[k]:bounds[k + 1]].astype(np.int32) for k, u in enumerate(uniq)}


def extract_submatrix(A, vertices: np.ndarray) -> Tuple[sp.csr_matrix, np.ndarray]:
    """_components.pyx:180-251: the subgraph induced by ``vertices`` (sorted),
    renumbered 0..len-1, and the new -> old mapping."""
    A = _csr(A)
    verts = np.asarray(vertices, dtype=np.int32)
    old_to_new = np.full(A.shape[0], -1, dtype=np.int64)
    old_to_new[verts] = np.arange(verts.size)
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    keep = (old_to_new[rows] >= 0) & (old_to_new[A.indices] >= 0)
    sub = sp.csr_matrix((A.data[keep].astype(np.int32), (old_to_new[rows[keep]], old_to_new[A.indices[keep]])),
                        shape=(verts.size, verts.size))
    sub.sort_indices()
    return sub, verts


def merge_component_matrices(components: List[Tuple[sp.csr_matrix, np.ndarray]], n_total: int) -> sp.csr_matrix:
    """_components.pyx:254-323: place component submatrices back at their
    original vertex indices."""
    rows, cols, vals = [], [], []
    for sub, mapping in components:
        c = sp.coo_matrix(sub)
        m = np.asarray(mapping, dtype=np.int64)
        rows.append(m[c.row])
        cols.append(m[c.col])
        vals.append(c.data.astype(np.int32))
    if not rows:
        return sp.csr_matrix((n_total, n_total), dtype=np.int32)
    out = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                        shape=(n_total, n_total))
    out.sort_indices()
    return out


def component_wise_closure(A, method: str = "auto", densify_threshold: int = 500):
    """transitive_closure.py:85-195: close every connected component on its
    own and merge the closed components into one CSR matrix.

    The components are found on the GPU and closed together by the batched
    dense launcher (the reference closes them one at a time, by Floyd-Warshall
    below ``densify_threshold`` vertices and by its Dijkstra closure above;
    in this build both give the Floyd-Warshall closure, so ``method`` and the
    threshold do not change the result).  The diameter is the largest
    component diameter.  A component with a cycle raises CycleError (vertex
    numbered within that component, as the reference's per-component call
    reports it), the first such component in component order.
    """
    from ...batched import BatchedClosure
    from ...core import raise_fw_errors
    from ..._dtypes import elem_type
    from ...types import TransitiveClosure
    from .transitive_closure import transitive_closure
    t = _lib.torch()
    A = _csr(A)
    n = A.shape[0]
    if n == 0:
        return TransitiveClosure(sp.csr_matrix((0, 0), dtype=np.int32), 0)
    A = sp.csr_matrix((A.data.astype(np.int32), A.indices, A.indptr), shape=A.shape)
    A.sum_duplicates()
    A.sort_indices()
    roots = _device_labels(A).cpu().numpy()
    ids = _ids_from_labels(roots, None).astype(np.int64)
    ncomp = int(ids.max()) + 1
    if ncomp == 1:
        res = transitive_closure(A, method="FW" if method == "FW" else "D")
        W = res.W if sp.isspmatrix(res.W) else sp.csr_matrix(np.asarray(res.W))
        return TransitiveClosure(W, res.diameter)
    # vertices of each component in increasing order, and each vertex's local index
    order = np.argsort(ids, kind="stable")
    sizes = np.bincount(ids, minlength=ncomp)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    local = np.empty(n, dtype=np.int64)
    local[order] = np.arange(n) - starts[ids[order]]
    b = BatchedClosure(sizes)
    try:
        dev = b.pool.device
        rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(A.indptr))
        comp = ids[rows]
        flat = (b.layout.offset[comp] + local[rows] * b.layout.npad[comp] + local[A.indices])
        with t.cuda.stream(b.stream):
            b.pool.zero_()
            b.pool.index_put_((t.from_numpy(flat).to(dev),), t.from_numpy(A.data.astype(np.int32)).to(dev))
        b.run(graph=False)
        verts = [order[starts[c]:starts[c] + sizes[c]] for c in range(ncomp)]

        def source(c):
            sub, _ = extract_submatrix(A, verts[c])
            return _lib.to_device(sub.toarray().astype(np.int32))
        outs = b.outcomes(source)
        et = elem_type(np.int32)
        for c, out in enumerate(outs):
            raise_fw_errors(out, et, "csgraph")
        diameter = max(out.diameter for out in outs)
        # merge on the device: every nonzero of the pool back to global indices
        nz = t.nonzero(b.pool).flatten()
        vals = b.pool[nz]
        offs = t.from_numpy(b.layout.offset).to(dev)
        npad = t.from_numpy(b.layout.npad).to(dev)
        c_of = t.searchsorted(offs, nz, right=True) - 1
        rel = nz - offs[c_of]
        lr, lc = rel // npad[c_of], rel % npad[c_of]
        gbase = t.from_numpy(starts).to(dev)
        gorder = t.from_numpy(order).to(dev)
        gr = gorder[gbase[c_of] + lr]
        gc = gorder[gbase[c_of] + lc]
        key = gr * n + gc
        srt = t.argsort(key)
        gr, gc, vals = gr[srt], gc[srt], vals[srt]
        indptr = t.zeros(n + 1, dtype=t.int64, device=dev)
        indptr[1:] = t.cumsum(t.bincount(gr, minlength=n), 0)
        W = sp.csr_matrix((vals.cpu().numpy(), gc.to(t.int32).cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
        return TransitiveClosure(W, int(diameter))
    finally:
        b.close()
