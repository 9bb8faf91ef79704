"""Canonical ordering of a closed graph and vertex permutation
(sparse/csgraph/_ordering.pyx:17-180 and _permutation.pyx:11-136 of the
reference; reference/ordering.py:13-53 states the same order).

Vertices are ranked on the device: rbg_order_metrics_* gives each vertex's
ancestor count (the sum of MSB over its row's nonzero entries), its max
relationship (its column's largest entry, at least 0) and its diagonal;
rbg_order_sort sorts the vertices on
    (-component size, component id, -ancestor count, max relationship,
     -diagonal, vertex id)
and rbg_permute_dense_* / rbg_permute_csr_* build P A P^-1 (row i of the
result is row p[i] of A, column j is column p[j]).  Component ids come from
the device component labels (find_components), numbered by each
component's smallest vertex as in the reference.
"""
from __future__ import annotations

import ctypes
from typing import Dict

import numpy as np
import scipy.sparse as sp

from ... import _lib
from ...types import Ordering
from ._components import find_components


def _value_dtype(a) -> np.dtype:
    return np.dtype(np.int64) if np.dtype(a.dtype) == np.dtype(np.int64) else np.dtype(np.int32)


def _sfx(dt: np.dtype) -> str:
    return "i64" if dt == np.dtype(np.int64) else "i32"


def _dense_device(A):
    A = np.asarray(A)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"expected a square matrix, got shape {A.shape}")
    dt = _value_dtype(A)
    return _lib.to_device(np.ascontiguousarray(A, dtype=dt)), dt


def _csr_device(A: sp.csr_matrix):
    dt = _value_dtype(A.data)
    return (_lib.to_device(np.ascontiguousarray(A.indptr, dtype=np.int64)),
            _lib.to_device(np.ascontiguousarray(A.indices, dtype=np.int32)),
            _lib.to_device(np.ascontiguousarray(A.data, dtype=dt)), dt)


def _sort(n: int, q: Dict[int, int], ids: np.ndarray, anc, maxrel, diag):
    """rbg_order_sort: the vertices in canonical order (device int32)."""
    t = _lib.torch()
    dev = _lib.device()
    ids = np.asarray(ids, dtype=np.int64)
    sizes = np.array([q[k] for k in range(len(q))], dtype=np.int32)[ids] if n else np.zeros(0, np.int32)
    comp = _lib.to_device(ids.astype(np.int32))
    csize = _lib.to_device(sizes)
    perm = t.empty(n, dtype=t.int32, device=dev)
    temp = t.empty(max(1, int(_lib.load().rbg_order_sort_temp_bytes(n))), dtype=t.uint8, device=dev)
    _lib.call("rbg_order_sort", n, _lib.ptr(comp), _lib.ptr(csize), _lib.ptr(anc), _lib.ptr(maxrel), _lib.ptr(diag),
              _lib.ptr(perm), _lib.ptr(temp), temp.numel(), _lib.stream_handle())
    return perm


def _metrics_dense(dA, n: int, dt):
    t = _lib.torch()
    dev = _lib.device()
    anc, maxrel, diag = (t.empty(n, dtype=t.int64, device=dev) for _ in range(3))
    _lib.call(f"rbg_order_metrics_dense_{_sfx(dt)}", _lib.ptr(dA), n, _lib.ptr(anc), _lib.ptr(maxrel), _lib.ptr(diag),
              _lib.stream_handle())
    return anc, maxrel, diag


def _metrics_csr(indptr, indices, data, n: int, dt):
    t = _lib.torch()
    dev = _lib.device()
    anc, maxrel, diag = (t.empty(n, dtype=t.int64, device=dev) for _ in range(3))
    _lib.call(f"rbg_order_metrics_csr_{_sfx(dt)}", _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(data), n,
              _lib.ptr(anc), _lib.ptr(maxrel), _lib.ptr(diag), _lib.stream_handle())
    return anc, maxrel, diag


def _get_permutation(A, q: Dict[int, int], ids) -> np.ndarray:
    """_ordering.pyx:17-47: the canonical vertex order of a dense closed matrix."""
    dA, dt = _dense_device(A)
    n = dA.shape[0]
    if n == 0:
        return np.zeros(0, dtype=np.int32)
    return _sort(n, q, ids, *_metrics_dense(dA, n, dt)).cpu().numpy()


def _get_permutation_sparse(A, q: Dict[int, int], ids) -> np.ndarray:
    """_ordering.pyx:50-96: the same order from CSR rows and columns."""
    A = A.tocsr() if sp.isspmatrix(A) else sp.csr_matrix(np.asarray(A))
    n = A.shape[0]
    if n == 0:
        return np.zeros(0, dtype=np.int32)
    indptr, indices, data, dt = _csr_device(A)
    return _sort(n, q, ids, *_metrics_csr(indptr, indices, data, n, dt)).cpu().numpy()


def _checked_perm(p, n: int):
    """The device copy of a caller's permutation, after checking on the host
    that it names every vertex once (the kernels index with it)."""
    p = np.ascontiguousarray(p, dtype=np.int32)
    if p.shape != (n,) or not np.array_equal(np.sort(p), np.arange(n, dtype=np.int32)):
        raise ValueError(f"p must be a permutation of range({n})")
    return _lib.to_device(p)


def _permute_dense_device(dA, n: int, dt, perm, upper: bool):
    t = _lib.torch()
    B = t.empty_like(dA)
    _lib.call(f"rbg_permute_dense_{_sfx(dt)}", _lib.ptr(dA), n, _lib.ptr(perm), _lib.ptr(B), int(upper),
              _lib.stream_handle())
    return B


def _permute_csr_device(indptr, indices, data, n: int, nnz: int, dt, perm, stream=None):
    t = _lib.torch()
    dev = _lib.device()
    L = _lib.load()
    sfx = _sfx(dt)
    ws = t.empty(int(getattr(L, f"rbg_permute_csr_workspace_bytes_{sfx}")(n, nnz)), dtype=t.uint8, device=dev)
    optr = t.empty(n + 1, dtype=t.int64, device=dev)
    oidx = t.empty(max(nnz, 1), dtype=t.int32, device=dev)
    oval = t.empty(max(nnz, 1), dtype=data.dtype, device=dev)
    _lib.call(f"rbg_permute_csr_{sfx}", _lib.ptr(indptr), _lib.ptr(indices), _lib.ptr(data), n, nnz, _lib.ptr(perm),
              _lib.ptr(optr), _lib.ptr(oidx), _lib.ptr(oval), _lib.ptr(ws), ws.numel(), _lib.stream_handle(stream))
    return optr, oidx[:nnz], oval[:nnz]


def _upper_csr(optr, oidx, oval, n: int):
    """Keeps the entries on or above the diagonal (_permutation.pyx:138-157):
    rbg_csr_upper, counting pass then filling pass."""
    t = _lib.torch()
    from ...sparse_dev import glue_temp
    dev = optr.device
    fn = f"rbg_csr_upper_{'i32' if oval.dtype == t.int32 else 'i64'}"
    ptr = t.empty(n + 1, dtype=t.int64, device=dev)
    temp = glue_temp(n, dev)
    sh = _lib.stream_handle()
    _lib.call(fn, _lib.ptr(optr), _lib.ptr(oidx), _lib.ptr(oval), n, _lib.ptr(ptr), None, None, _lib.ptr(temp),
              temp.numel(), sh)
    nnz = int(ptr[n].item())
    oi = t.empty(max(nnz, 1), dtype=t.int32, device=dev)
    ov = t.empty(max(nnz, 1), dtype=oval.dtype, device=dev)
    _lib.call(fn, _lib.ptr(optr), _lib.ptr(oidx), _lib.ptr(oval), n, _lib.ptr(ptr), _lib.ptr(oi), _lib.ptr(ov),
              _lib.ptr(temp), temp.numel(), sh)
    return ptr, oi[:nnz], ov[:nnz]


def permute_sparse(A, p, assume_upper_triangular: bool = False) -> sp.csr_matrix:
    """_permutation.pyx:67-136: P A P^-1 of a sparse matrix, rows sorted."""
    A = A.tocsr() if sp.isspmatrix(A) else sp.csr_matrix(np.asarray(A))
    n = A.shape[0]
    if n == 0:
        return sp.csr_matrix(A.shape, dtype=A.dtype)
    indptr, indices, data, dt = _csr_device(A)
    perm = _checked_perm(p, n)
    optr, oidx, oval = _permute_csr_device(indptr, indices, data, n, int(A.nnz), dt, perm)
    if assume_upper_triangular:
        optr, oidx, oval = _upper_csr(optr, oidx, oval, n)
    return sp.csr_matrix((oval.cpu().numpy(), oidx.cpu().numpy(), optr.cpu().numpy()), shape=A.shape)


def permute(A, p, assume_upper_triangular: bool = False):
    """_permutation.pyx:11-66: P A P^-1 (B[i][j] = A[p[i]][p[j]]), dense or sparse."""
    if sp.isspmatrix(A):
        return permute_sparse(A, p, assume_upper_triangular)
    dA, dt = _dense_device(A)
    n = dA.shape[0]
    if n == 0:
        return np.zeros((0, 0), dtype=dt)
    perm = _checked_perm(p, n)
    return _permute_dense_device(dA, n, dt, perm, assume_upper_triangular).cpu().numpy()


def avos_canonical_ordering(A) -> Ordering:
    """_ordering.pyx:160-180: components in adjacent rows, largest first (then
    by component id); within a component by ancestor count descending, max
    relationship ascending, color descending and vertex id.  A is assumed to
    be transitively closed.  Dense input gives a dense result, sparse input a
    CSR result."""
    q: Dict[int, int] = {}
    ids = find_components(A, q)
    if sp.isspmatrix(A):
        A = A.tocsr()
        n = A.shape[0]
        if n == 0:
            return Ordering(sp.csr_matrix(A.shape, dtype=A.dtype), np.zeros(0, dtype=np.int32), q)
        indptr, indices, data, dt = _csr_device(A)
        perm = _sort(n, q, ids, *_metrics_csr(indptr, indices, data, n, dt))
        optr, oidx, oval = _permute_csr_device(indptr, indices, data, n, int(A.nnz), dt, perm)
        B = sp.csr_matrix((oval.cpu().numpy(), oidx.cpu().numpy(), optr.cpu().numpy()), shape=A.shape)
        return Ordering(B, perm.cpu().numpy(), q)
    dA, dt = _dense_device(A)
    n = dA.shape[0]
    if n == 0:
        return Ordering(np.zeros((0, 0), dtype=dt), np.zeros(0, dtype=np.int32), q)
    perm = _sort(n, q, ids, *_metrics_dense(dA, n, dt))
    B = _permute_dense_device(dA, n, dt, perm, False)
    return Ordering(B.cpu().numpy(), perm.cpu().numpy(), q)


def canonical_order_device(C, stream=None):
    """The canonical ordering of a closed graph held as a device CSR
    (sparse_dev.DevCSR), without leaving the device: component labels
    (rbg_components), ids by smallest vertex and sizes, the ordering metrics,
    the sort and the CSR permutation.  Returns (permutation, permuted DevCSR,
    component sizes by id), all device tensors."""
    from ... import sparse_dev as sd
    t = _lib.torch()
    dev = C.indices.device
    n = C.n_rows
    sh = _lib.stream_handle(stream)
    dt = np.dtype(np.int64) if C.data.dtype == t.int64 else np.dtype(np.int32)
    labels = t.empty(max(n, 1), dtype=t.int32, device=dev)
    flags = t.zeros(2, dtype=t.int32, device=dev)
    rounds = ctypes.c_int32(0)
    _lib.call("rbg_components", _lib.ptr(C.indptr), _lib.ptr(C.indices), n, _lib.ptr(labels), _lib.ptr(flags),
              ctypes.byref(rounds), sh)
    # component ids by smallest vertex, and each vertex's component size
    comp, local, members, sizes, csize = (t.empty(n, dtype=t.int32, device=dev) for _ in range(5))
    starts = t.empty(n + 1, dtype=t.int64, device=dev)
    ncomp = t.empty(1, dtype=t.int64, device=dev)
    gtemp = sd.glue_temp(n, dev)
    _lib.call("rbg_components_layout", _lib.ptr(labels), n, _lib.ptr(comp), _lib.ptr(local), _lib.ptr(members),
              _lib.ptr(sizes), _lib.ptr(starts), _lib.ptr(ncomp), _lib.ptr(csize), _lib.ptr(gtemp), gtemp.numel(), sh)
    counts = sizes[:int(ncomp.item())]
    anc, maxrel, diag = (t.empty(n, dtype=t.int64, device=dev) for _ in range(3))
    _lib.call(f"rbg_order_metrics_csr_{_sfx(dt)}", _lib.ptr(C.indptr), _lib.ptr(C.indices), _lib.ptr(C.data), n,
              _lib.ptr(anc), _lib.ptr(maxrel), _lib.ptr(diag), sh)
    perm = t.empty(n, dtype=t.int32, device=dev)
    temp = t.empty(max(1, int(_lib.load().rbg_order_sort_temp_bytes(n))), dtype=t.uint8, device=dev)
    _lib.call("rbg_order_sort", n, _lib.ptr(comp), _lib.ptr(csize), _lib.ptr(anc), _lib.ptr(maxrel), _lib.ptr(diag),
              _lib.ptr(perm), _lib.ptr(temp), temp.numel(), sh)
    optr, oidx, oval = _permute_csr_device(C.indptr, C.indices, C.data, n, int(C.nnz), dt, perm, stream)
    return perm, sd.DevCSR(optr, oidx, oval, C.shape), counts
