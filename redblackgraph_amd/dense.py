"""Device-level dense operations: Floyd-Warshall closure, AVOS product,
relational composition.  Inputs and outputs are torch tensors resident on the
GPU, in the widened device domain (see _dtypes).  The reference-shaped API in
core.py wraps these.
"""
from __future__ import annotations

import threading
from dataclasses import dataclass

from . import _lib
from ._dtypes import ElemType, msb

FW_TILE = 128
_tls = threading.local()


def _workspace(npad: int, elem: int):
    """The calling thread's snapshot workspace (one per thread and device:
    threads run closures on their own streams and must not share it)."""
    t = _lib.torch()
    dev = _lib.device()
    cache = getattr(_tls, "ws_cache", None)
    if cache is None:
        cache = _tls.ws_cache = {}
    key = (dev.index, npad, elem)
    ws = cache.get(key)
    if ws is None:
        nbytes = int(_lib.load().rbg_fw_tri_workspace_bytes(npad, elem))
        ws = t.empty(nbytes, dtype=t.uint8, device=dev)
        cache.clear()
        cache[key] = ws
    return ws


def padded_order(n: int) -> int:
    return max(FW_TILE, ((n + FW_TILE - 1) // FW_TILE) * FW_TILE)


@dataclass
class FwOutcome:
    status: _lib.RbgStatus
    npad: int
    mode: str  # "fast" | "general"
    # "full" | "lower" | "upper" (triangular schedules; "lower" is the
    # persistent closure) | "lower-steps" (one launch per pivot step)
    schedule: str = "full"

    @property
    def diameter(self) -> int:
        return msb(self.status.diameter_max)

    def key_kij(self, key: int):
        n = self.npad
        return key // (n * n), (key // n) % n, key % n


def _run(work, et: ElemType, mode: int, st: _lib.Status, stream, tri: int = 0):
    t = _lib.torch()
    npad = work.shape[0]
    elem = work.element_size()
    ws = _workspace(npad, elem)
    fn = "rbg_fw_dense_tri_i32" if work.dtype == t.int32 else "rbg_fw_dense_tri_i64"
    _lib.call(fn, _lib.ptr(work), npad, npad, et.width, int(et.unsigned), mode, tri,
              _lib.ptr(ws), ws.numel(), st.p, _lib.stream_handle(stream))


def status() -> _lib.Status:
    """A per-thread reusable status block (device + pinned host mirror)."""
    st = getattr(_tls, "status", None)
    if st is None:
        st = _lib.Status()
        _tls.status = st
    return st


_TRI = {"full": 0, "lower": 1, "upper": 2, "lower-steps": 3}


def fw_closure(W, et: ElemType, stream=None, force_general: bool = False, schedule: str = "auto"):
    """Close an n x n device matrix.  Returns (closed n x n tensor, FwOutcome).

    The input tensor is not modified.  Runs the fast kernels when the input is
    inside their domain (rbg_fw_precheck) and falls back to the general kernels
    otherwise -- both on the GPU.  schedule "auto" takes the triangular
    schedule when the precheck finds every off-diagonal entry on one side of
    the diagonal (a pedigree numbered parents first is lower triangular), else
    the full one; "full" forces the full schedule (the results are the same).
    A lower triangular closure runs as one persistent launch; "steps" keeps
    one launch per pivot step there (the same results again).
    """
    if schedule not in ("auto", "full", "steps"):
        raise ValueError(f"unknown schedule {schedule!r}")
    t = _lib.torch()
    n = W.shape[0]
    npad = padded_order(n)

    def fresh():
        if npad == n:
            return W.clone()
        w = t.zeros((npad, npad), dtype=W.dtype, device=W.device)
        w[:n, :n] = W
        return w

    st = status()
    st.reset(stream)
    fast = et.native and not force_general
    sched = "full"
    if fast:
        # one pass: copy, fast-domain check and encoding (rbg_fw_load_*)
        src = W if (W.stride(1) == 1 and W.stride(0) >= n) else W.contiguous()
        work = t.empty((npad, npad), dtype=W.dtype, device=W.device)
        fn = "rbg_fw_load_i32" if work.dtype == t.int32 else "rbg_fw_load_i64"
        _lib.call(fn, _lib.ptr(src), n, src.stride(0), _lib.ptr(work), npad, st.p, _lib.stream_handle(stream))
        s = st.read(stream)
        if s.flags & _lib.F_PRECOND_FAIL:
            fast = False
            st.reset(stream)
        elif schedule in ("auto", "steps"):
            if not s.flags & _lib.F_HAS_UPPER:
                sched = "lower" if schedule == "auto" else "lower-steps"
            elif not s.flags & _lib.F_HAS_LOWER:
                sched = "upper"
    if not fast:
        work = fresh()
        sched = "full"
    _run(work, et, 2 if fast else 1, st, stream, _TRI[sched])
    s = st.read(stream)
    if fast and s.flags & _lib.F_STALLED:
        # a wait of the persistent closure ran out: the same schedule, one
        # launch per pivot step, on the original input
        work = fresh()
        st.reset(stream)
        sched = "lower-steps"
        _run(work, et, 0, st, stream, _TRI[sched])
        s = st.read(stream)
    mode = "fast" if fast else "general"
    # A cycle makes a diagonal value >= 2, after which products through the
    # pivot cell (skipped by the fast kernels) can raise events of their own:
    # the error path reruns in general mode, which records every event.
    if fast and (s.flags & _lib.F_NEED_GENERAL or s.cycle_count):
        work = fresh()
        st.reset(stream)
        _run(work, et, 1, st, stream)
        s = st.read(stream)
        mode = "general"
        sched = "full"
    return work[:n, :n], FwOutcome(s, npad, mode, sched)


# Products with fewer multiply-adds than this take the general kernel directly
# (the fast path's precheck and operand encoding cost more than they save).
MATMUL_FAST_MIN_TERMS = 1 << 24


def matmul(A, B, et: ElemType, stream=None, force_general: bool = False):
    """C = A (x) B for device matrices (M x K) and (K x N).  Native-width
    operands inside the fast domain (rbg_matmul_precheck: closures of
    pedigrees) take rbg_matmul_fast_*; others, and a fast run whose MSB guard
    fails, take the general kernel with the full arithmetic and events."""
    t = _lib.torch()
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    C = t.empty((M, N), dtype=A.dtype, device=A.device)
    st = status()
    sfx = "i32" if A.dtype == t.int32 else "i64"
    sh = _lib.stream_handle(stream)
    A = A.contiguous()
    B = B.contiguous()
    if et.native and not force_general and M * K * N >= MATMUL_FAST_MIN_TERMS:
        st.reset(stream)
        _lib.call(f"rbg_matmul_precheck_{sfx}", _lib.ptr(A), K, M, K, st.p, sh)
        _lib.call(f"rbg_matmul_precheck_{sfx}", _lib.ptr(B), N, K, N, st.p, sh)
        if not st.read(stream).flags & _lib.F_PRECOND_FAIL:
            ws = t.empty(int(_lib.load().rbg_matmul_fast_workspace_bytes(M, K, N, A.element_size())),
                         dtype=t.uint8, device=A.device)
            st.reset(stream)
            _lib.call(f"rbg_matmul_fast_{sfx}", _lib.ptr(A), K, _lib.ptr(B), N, _lib.ptr(C), N, M, K, N,
                      _lib.ptr(ws), ws.numel(), st.p, sh)
            s = st.read(stream)
            if not s.flags & _lib.F_NEED_GENERAL:
                return C, s
    st.reset(stream)
    _lib.call(f"rbg_matmul_{sfx}", _lib.ptr(A), K, _lib.ptr(B), N, _lib.ptr(C), N, M, K, N, et.width,
              int(et.unsigned), st.p, sh)
    return C, st.read(stream)


def edge_compose(R, alpha: int, beta: int, rel: int, et: ElemType, stream=None):
    t = _lib.torch()
    n = R.shape[0]
    out = t.empty_like(R)
    st = status()
    st.reset(stream)
    fn = "rbg_edge_compose_i32" if R.dtype == t.int32 else "rbg_edge_compose_i64"
    _lib.call(fn, _lib.ptr(R.contiguous()), n, alpha, beta, int(rel), _lib.ptr(out), et.width,
              int(et.unsigned), st.p, _lib.stream_handle(stream))
    return out, st.read(stream)


def edge_compose_many(R, edges, et: ElemType, stream=None, inplace: bool = False):
    """Apply edge compositions one after another without host round trips:
    each edge reads the previous result.  The compositions run in place on
    one copy of R (R itself too when ``inplace``), touching only the rows
    that reach each edge's source.  Returns (closure, status); status.flags
    has RBG_F_CYCLE when some edge would have closed a cycle, and the result
    is then meaningless."""
    t = _lib.torch()
    n = R.shape[0]
    st = status()
    st.reset(stream)
    fn = "rbg_edge_compose_inplace_i32" if R.dtype == t.int32 else "rbg_edge_compose_inplace_i64"
    sh = _lib.stream_handle(stream)
    cur = R if (inplace and R.is_contiguous()) else R.contiguous().clone()
    scratch = t.empty(int(_lib.load().rbg_edge_compose_scratch_bytes(n)), dtype=t.uint8, device=R.device)
    for a, b, rel in edges:
        _lib.call(fn, _lib.ptr(cur), n, int(a), int(b), int(rel), _lib.ptr(scratch), et.width, int(et.unsigned),
                  st.p, sh)
    return cur, st.read(stream)


def vertex_compose(u, R, v, color: int, et: ElemType, stream=None):
    t = _lib.torch()
    n = R.shape[0]
    out = t.zeros((n + 1, n + 1), dtype=R.dtype, device=R.device)
    st = status()
    st.reset(stream)
    fn = "rbg_vertex_compose_i32" if R.dtype == t.int32 else "rbg_vertex_compose_i64"
    _lib.call(fn, _lib.ptr(u.contiguous()), _lib.ptr(R.contiguous()), _lib.ptr(v.contiguous()), int(color), n,
              _lib.ptr(out), et.width, int(et.unsigned), st.p, _lib.stream_handle(stream))
    return out, st.read(stream)
