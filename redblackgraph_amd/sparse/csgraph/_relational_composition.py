"""Relational composition of the csgraph family
(sparse/csgraph/_relational_composition.pyx:37-113 of the reference), on the
device's composition kernels.

These are the Cython versions: int32 arithmetic (DTYPE) without the C core's
checks, so a composition that would close a cycle or overflow is computed
and returned, not raised (the core versions, core/redblack.py:52-97, raise
ValueError on a cycle).  The kernels compute the same values as the C core's
composition; the flags they raise are ignored here.
"""
from __future__ import annotations

import numpy as np

from ... import dense, _lib
from ..._dtypes import elem_type


def vertex_relational_composition(u, R, v, color):
    """_relational_composition.pyx:37-65: the (N+1) x (N+1) closure of R with
    a new vertex lambda whose row is u (lambda's direct relationships) and
    column v (the vertices related to lambda), of colour color (-1 / 1):
        row N    = u (x) R,  column N = R (x) v,
        [i][j]   = R[i][j] (+) (R (x) v)[i] (x) (u (x) R)[j],  [N][N] = color.
    Returns a numpy int32 array."""
    R = np.ascontiguousarray(np.asarray(R), dtype=np.int32)
    n = R.shape[0]
    assert n == R.shape[1]
    uu = np.ascontiguousarray(np.asarray(u, dtype=np.int32).reshape((n,)))
    vv = np.ascontiguousarray(np.asarray(v, dtype=np.int32).reshape((n,)))
    et = elem_type(np.int32)
    out, _ = dense.vertex_compose(_lib.to_device(uu), _lib.to_device(R), _lib.to_device(vv), int(color), et)
    return out.cpu().numpy()


def edge_relational_composition(R, alpha, beta, relationship):
    """_relational_composition.pyx:93-113: the closure R with the edge
    alpha -> beta of value `relationship` composed in: row alpha becomes
    u = (R with R[alpha][beta] = relationship)[alpha] (x) R, and every other
    row i takes R[i][j] (+) R[i][alpha] (x) u[j].  R is not modified.
    Returns a numpy int32 array."""
    R = np.ascontiguousarray(np.asarray(R), dtype=np.int32)
    n = R.shape[0]
    assert n == R.shape[1]
    et = elem_type(np.int32)
    out, _ = dense.edge_compose(_lib.to_device(R), int(alpha), int(beta), int(relationship), et)
    return out.cpu().numpy()
