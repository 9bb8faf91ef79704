"""Dense AVOS API, shaped like redblackgraph.core (core/redblack.py:1-130,
core/avos.py:1-20).

``array`` is an ndarray subclass whose ``@`` is the AVOS product and whose
``transitive_closure()`` is the Floyd-Warshall closure; both run on the GPU
through librbg_hip.so.  Host arrays are copied to the device and back; torch
tensors already on the device can use redblackgraph_amd.dense directly.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, dense
from ._dtypes import as_unsigned_text, elem_type, narrow, widen
from .types import TransitiveClosure

__all__ = ["array", "transitive_closure", "vertex_relational_composition", "edge_relational_composition",
           "avos_sum", "avos_product", "einsum"]


# ---------------------------------------------------------------------------
# scalar arithmetic (redblackgraphmodule.c:490-534: 64-bit unsigned rules)
# ---------------------------------------------------------------------------
def _as_longlong(v) -> int:
    v = int(v)
    if not (-(1 << 63) <= v < (1 << 63)):
        raise OverflowError("Python int too large to convert to C long")
    return v


def avos_sum(lhs: int, rhs: int) -> int:
    a, b = _as_longlong(lhs), _as_longlong(rhs)
    r = int(_lib.load().rbg_host_avos_sum_i64(a, b)) & ((1 << 64) - 1)
    return -1 if r == (1 << 64) - 1 else r


def avos_product(lhs: int, rhs: int) -> int:
    a, b = _as_longlong(lhs), _as_longlong(rhs)
    ev = ctypes.c_uint32(0)
    r = int(_lib.load().rbg_host_avos_product_u64(a, b, ctypes.byref(ev)))
    if ev.value & _lib.EV_OVERFLOW_WIDTH:
        bits = (a & ((1 << 64) - 1) if a != -1 else 1).bit_length() - 1 + (b & ((1 << 64) - 1)).bit_length() - 1
        raise OverflowError(f"Avos product of {a & ((1 << 64) - 1)} and {b & ((1 << 64) - 1)}, results in an "
                            f"overflow. (Result size would require {bits + 1} bits; Type provides 64 bits)")
    if ev.value & _lib.EV_OVERFLOW_RED:
        raise OverflowError(f"Avos product of {a & ((1 << 64) - 1)} and {b & ((1 << 64) - 1)}, results in an "
                            f"overflow. Result of avos product collides with 'red' 1 (-1).")
    return -1 if r == (1 << 64) - 1 else r


# ---------------------------------------------------------------------------
# error mapping for the dense closure
# ---------------------------------------------------------------------------
def _overflow_message(a: int, b: int, et) -> str:
    ua, ub = as_unsigned_text(a, 64), as_unsigned_text(b, 64)
    return f"Avos product of {ua} and {ub}, results in an overflow."


def raise_fw_errors(out: dense.FwOutcome, et, style: str = "core"):
    """Raise what the reference raises after its FW loop.

    style "core": warshall.c.in:82-87 -- the loop runs to the end and the error
    set last survives (ValueError for a cycle, OverflowError for overflow).
    style "csgraph": _shortest_path.pyx:374-386 -- CycleError at the first
    detection.
    """
    s = out.status
    if s.events & _lib.EV_INVALID:
        raise ValueError("Invalid input. Negative values (aside from -1) are not allowed.")
    has_cycle = s.cycle_count > 0
    has_ov = s.ov_first_key != _lib.NO_KEY and (s.events & (_lib.EV_OVERFLOW_WIDTH | _lib.EV_OVERFLOW_RED | _lib.EV_SIGN))
    if style == "csgraph":
        if has_cycle:
            from .sparse.csgraph.cycleerror import CycleError
            k, i, _ = out.key_kij(s.cycle_first_key)
            raise CycleError(f"Error: cycle detected! Vertex {i} has a path to itself. "
                             f"A({i},{k})={s.cycle_first_a}, A({k},{i})={s.cycle_first_b}", vertex=int(i))
        if has_ov:
            raise OverflowError(_overflow_message(s.ov_last_a, s.ov_last_b, et))
        return
    if not (has_cycle or has_ov):
        return
    cyc_key = (s.cycle_last_key * 2 + 1) if has_cycle else -1
    ov_key = (s.ov_last_key * 2) if has_ov else -1
    if cyc_key > ov_key:
        k, i, _ = out.key_kij(s.cycle_last_key)
        a = as_unsigned_text(s.cycle_last_a, et.width)
        b = as_unsigned_text(s.cycle_last_b, et.width)
        raise ValueError(f"Error: cycle detected! Vertex {i} has a path to itself. A({i},{k})={a}, A({k},{i})={b}")
    raise OverflowError(_overflow_message(s.ov_last_a, s.ov_last_b, et))


def _check_events(st, et):
    if st.events & _lib.EV_INVALID:
        raise ValueError("Invalid input. Negative values (aside from -1) are not allowed.")
    if st.events & (_lib.EV_OVERFLOW_WIDTH | _lib.EV_OVERFLOW_RED | _lib.EV_SIGN):
        raise OverflowError(f"Avos product results in an overflow for element type {et.dtype}.")


# ---------------------------------------------------------------------------
# host <-> device helpers
# ---------------------------------------------------------------------------
def _dev(a: np.ndarray, et):
    return _lib.to_device(widen(np.ascontiguousarray(a), et))


def _host(t, et) -> np.ndarray:
    return narrow(t.cpu().numpy(), et)


def warshall(A: np.ndarray):
    """(W, diameter) of the dense closure, raising like the C core."""
    A = np.asarray(A)
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise ValueError(f"warshall expects a square matrix, got shape {A.shape}")
    et = elem_type(A.dtype)
    if A.shape[0] == 0:
        return A.copy(), 0
    W, out = dense.fw_closure(_dev(A, et), et)
    raise_fw_errors(out, et, "core")
    return _host(W, et), out.diameter


def _matmul2d(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    dt = np.result_type(A.dtype, B.dtype)
    et = elem_type(dt)
    A = np.asarray(A, dtype=dt)
    B = np.asarray(B, dtype=dt)
    if A.shape[1] != B.shape[0]:
        raise ValueError(f"operands could not be broadcast together with shapes {A.shape} {B.shape}")
    if A.size == 0 or B.size == 0:
        return np.zeros((A.shape[0], B.shape[1]), dtype=dt)
    C, st = dense.matmul(_dev(A, et), _dev(B, et), et)
    _check_events(st, et)
    return _host(C, et)


def einsum(subscripts, a, b, avos=True):
    """The four contractions the reference's _Avos.__matmul__ uses
    (core/redblack.py:14-38): 'ij,jk', 'i,...ij', '...i,i', 'i,i'."""
    if not avos:
        raise ValueError("Avos should be specified in all cases.")
    a = np.asarray(a)
    b = np.asarray(b)
    if subscripts == "ij,jk":
        return _matmul2d(a, b)
    if subscripts == "i,...ij":
        return _matmul2d(a.reshape(1, -1), b.reshape(b.shape[-2], b.shape[-1])).reshape(-1)
    if subscripts == "...i,i":
        return _matmul2d(a.reshape(-1, a.shape[-1]), b.reshape(-1, 1)).reshape(a.shape[:-1])
    if subscripts == "i,i":
        return _matmul2d(a.reshape(1, -1), b.reshape(-1, 1)).reshape(())[()]
    raise ValueError(f"unsupported AVOS contraction {subscripts!r}")


class _Avos(np.ndarray):
    def __matmul__(self, other):
        other = np.asarray(other)
        if self.ndim == 1:
            if other.ndim == 1:
                return einsum("i,i", self, other)
            if other.ndim == 2:
                return einsum("i,...ij", self, other)
        if self.ndim == 2:
            if other.ndim == 1:
                return einsum("...i,i", self, other)
            if other.ndim == 2:
                return einsum("ij,jk", self, other)
        raise ValueError(f"Unexpected dimensionality. self: {self.ndim}, other: {other.ndim}")

    def __rmatmul__(self, other):
        other = np.asarray(other)
        if self.ndim == 1:
            if other.ndim == 1:
                return einsum("i,i", other, self)
            if other.ndim == 2:
                return einsum("...i,i", other, self)
        if self.ndim == 2:
            if other.ndim == 1:
                return einsum("i,...ij", other, self)
            if other.ndim == 2:
                return einsum("ij,jk", other, self)
        raise ValueError(f"Unexpected dimensionality. self: {self.ndim}, other: {other.ndim}")

    def cardinality(self):
        """Vertex counts by colour (core/redblack.py:40-47): the diagonal holds
        -1 for red and 1 for black, so with r + b = n and b - r = trace the
        black count is (n + trace) / 2."""
        n = self.shape[0]
        black = int((n + int(np.trace(self))) / 2)
        return {"red": n - black, "black": black}

    def transitive_closure(self) -> TransitiveClosure:
        W, d = warshall(np.asarray(self))
        return TransitiveClosure(W.view(type(self)), d)

    def vertex_relational_composition(self, u, v, c, compute_closure=False):
        R_star = warshall(np.asarray(self))[0] if compute_closure else np.asarray(self)
        n = R_star.shape[0]
        u = np.asarray(u).reshape(-1)
        v = np.asarray(v).reshape(-1)
        if u.shape[0] != n or v.shape[0] != n:
            raise ValueError("u and v must have one entry per vertex")
        et = elem_type(R_star.dtype)
        out, st = dense.vertex_compose(_dev(u.astype(R_star.dtype), et), _dev(R_star, et),
                                       _dev(v.astype(R_star.dtype), et), int(c), et)
        if st.flags & _lib.F_CYCLE:
            idx, ui, vi = st.aux[0], st.aux[1], st.aux[2]
            raise ValueError(f"Relational composition would result in a cycle. Idx: {idx}, u_i: {ui}, v_i: {vi}")
        _check_events(st, et)
        return _host(out, et).view(type(self))

    def edge_relational_composition(self, alpha, beta, relationship, compute_closure=False):
        R_star = warshall(np.asarray(self))[0] if compute_closure else np.asarray(self)
        if R_star[beta][alpha] != 0:
            raise ValueError("Relational composition would result in a cycle.")
        et = elem_type(R_star.dtype)
        out, st = dense.edge_compose(_dev(R_star, et), int(alpha), int(beta), int(relationship), et)
        _check_events(st, et)
        return _host(out, et).view(type(self))


class array(_Avos, np.ndarray):
    def __new__(cls, *args, **kwargs):
        if "dtype" in kwargs:
            dt = np.dtype(kwargs["dtype"])
            if dt.kind == "u":
                kw = {k: v for k, v in kwargs.items() if k != "dtype"}
                return np.asarray(*args, **kw).astype(dt).view(cls)
        return np.asarray(*args, **kwargs).view(cls)

    def __matmul__(self, other):
        r = super().__matmul__(other)
        return r.view(array) if isinstance(r, np.ndarray) else r

    def __rmatmul__(self, other):
        r = super().__rmatmul__(other)
        return r.view(array) if isinstance(r, np.ndarray) else r


def transitive_closure(A) -> TransitiveClosure:
    if not isinstance(A, _Avos):
        A = array(A)
    return A.transitive_closure()


def vertex_relational_composition(u, A, v, c, compute_closure=False):
    if not isinstance(A, _Avos):
        A = array(A)
    return A.vertex_relational_composition(u, v, c, compute_closure=compute_closure)


def edge_relational_composition(A, alpha, beta, relationship, compute_closure=False):
    if not isinstance(A, _Avos):
        A = array(A)
    return A.edge_relational_composition(alpha, beta, relationship, compute_closure=compute_closure)
