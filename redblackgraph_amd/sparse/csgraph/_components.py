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


def _device_labels_dev(dA):
    """_device_labels for a device CSR (sparse_dev.DevCSR)."""
    t = _lib.torch()
    n = dA.n_rows
    labels = t.empty(max(n, 1), dtype=t.int32, device=dA.indices.device)
    flags = t.zeros(2, dtype=t.int32, device=dA.indices.device)
    rounds = ctypes.c_int32(0)
    _lib.call("rbg_components", _lib.ptr(dA.indptr), _lib.ptr(dA.indices), n, _lib.ptr(labels), _lib.ptr(flags),
              ctypes.byref(rounds), _lib.stream_handle())
    return labels


def find_components_sparse(A, q: Optional[Dict[int, int]] = None) -> np.ndarray:
    """_components.pyx:86-177: component id per vertex (uint32), ids in order
    of each component's smallest vertex; q receives id -> vertex count."""
    from ... import sparse_dev as sd
    A = _csr(A)
    n = A.shape[0]
    if n == 0:
        return np.zeros(0, dtype=np.uint32)
    t = _lib.torch()
    labels = _device_labels(A)
    dev = labels.device
    comp, local, members, sizes = (t.empty(n, dtype=t.int32, device=dev) for _ in range(4))
    starts = t.empty(n + 1, dtype=t.int64, device=dev)
    ncomp = t.empty(1, dtype=t.int64, device=dev)
    temp = sd.glue_temp(n, dev)
    _lib.call("rbg_components_layout", _lib.ptr(labels), n, _lib.ptr(comp), _lib.ptr(local), _lib.ptr(members),
              _lib.ptr(sizes), _lib.ptr(starts), _lib.ptr(ncomp), None, _lib.ptr(temp), temp.numel(), _lib.stream_handle())
    if q is not None:
        for cid, c in enumerate(sizes[:int(ncomp.item())].cpu().numpy()):
            q[int(cid)] = int(c)
    return comp.cpu().numpy().astype(np.uint32)


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


def _triangularity(dA) -> int:
    """1 when every off-diagonal entry of the device CSR dA lies below the
    diagonal, 2 when above, else 0 (counted on the device).  Every component
    then is too: its vertices keep their relative order."""
    from ...dense import status
    st = status()
    st.reset()
    _lib.call("rbg_csr_triangles", _lib.ptr(dA.indptr), _lib.ptr(dA.indices), dA.n_rows, st.p, _lib.stream_handle())
    s = st.read()
    below, diag = int(s.aux[0]), int(s.aux[1])
    above = dA.nnz - below - diag
    return 1 if above == 0 else (2 if below == 0 else 0)


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
    from ... import sparse_dev as sd
    A = sp.csr_matrix((A.data.astype(np.int32), A.indices, A.indptr), shape=A.shape)
    # sorted rows, duplicates combined, zeros dropped: on the device
    dA = sd.canonicalize(sd.from_scipy(A, np.int32))
    dev = dA.indices.device
    sh = _lib.stream_handle()
    labels = _device_labels_dev(dA)
    comp, local, members, sizes = (t.empty(n, dtype=t.int32, device=dev) for _ in range(4))
    starts = t.empty(n + 1, dtype=t.int64, device=dev)
    ncomp_d = t.empty(1, dtype=t.int64, device=dev)
    temp = sd.glue_temp(n, dev)
    _lib.call("rbg_components_layout", _lib.ptr(labels), n, _lib.ptr(comp), _lib.ptr(local), _lib.ptr(members),
              _lib.ptr(sizes), _lib.ptr(starts), _lib.ptr(ncomp_d), None, _lib.ptr(temp), temp.numel(), sh)
    ncomp = int(ncomp_d.item())
    if ncomp == 1:
        res = transitive_closure(sd.to_scipy(dA), method="FW" if method == "FW" else "D")
        W = res.W if sp.isspmatrix(res.W) else sp.csr_matrix(np.asarray(res.W))
        return TransitiveClosure(W, res.diameter)
    sizes_h = sizes[:ncomp].cpu().numpy().astype(np.int64)
    b = BatchedClosure(sizes_h, tri=_triangularity(dA))
    try:
        off = t.from_numpy(b.layout.offset.astype(np.int64)).to(dev)
        npad = t.from_numpy(b.layout.npad.astype(np.int32)).to(dev)
        b.stream.wait_stream(t.cuda.current_stream())  # the layout, on the current stream
        bsh = _lib.stream_handle(b.stream)
        with t.cuda.stream(b.stream):
            b.pool.zero_()
        # extract_submatrix for every component at once: each entry into its
        # component's padded matrix in the batch pool
        _lib.call("rbg_components_scatter", _lib.ptr(dA.indptr), _lib.ptr(dA.indices), _lib.ptr(dA.data), n,
                  _lib.ptr(comp), _lib.ptr(local), _lib.ptr(off), _lib.ptr(npad), _lib.ptr(b.pool), bsh)
        b.run(graph=False)
        host = {}

        def source(c):
            # the component's input again (only for components the fast kernels hand back)
            if not host:
                host["members"] = members.cpu().numpy()
                host["starts"] = starts.cpu().numpy()
            s0 = int(host["starts"][c])
            sub, _ = extract_submatrix(A, np.sort(host["members"][s0:s0 + int(sizes_h[c])]))
            return _lib.to_device(sub.toarray().astype(np.int32))
        outs = b.outcomes(source)
        et = elem_type(np.int32)
        for c, out in enumerate(outs):
            raise_fw_errors(out, et, "csgraph")
        diameter = max(out.diameter for out in outs)
        # merge_component_matrices on the device: the closed components back
        # into one CSR in global vertex order
        indptr = t.empty(n + 1, dtype=t.int64, device=dev)
        args = [_lib.ptr(b.pool), _lib.ptr(off), _lib.ptr(npad), _lib.ptr(comp), _lib.ptr(local), _lib.ptr(members),
                _lib.ptr(starts), _lib.ptr(sizes), n, _lib.ptr(indptr)]
        _lib.call("rbg_components_merge", *args, None, None, _lib.ptr(temp), temp.numel(), bsh)
        b.stream.synchronize()
        nnz = int(indptr[n].item())
        oi = t.empty(max(nnz, 1), dtype=t.int32, device=dev)
        ov = t.empty(max(nnz, 1), dtype=t.int32, device=dev)
        _lib.call("rbg_components_merge", *args, _lib.ptr(oi), _lib.ptr(ov), _lib.ptr(temp), temp.numel(), bsh)
        b.stream.synchronize()
        W = sp.csr_matrix((ov[:nnz].cpu().numpy(), oi[:nnz].cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
        return TransitiveClosure(W, int(diameter))
    finally:
        b.close()
