"""Device-level sparse operations on CSR matrices held in torch tensors.

Device CSR convention: int64 row pointers, int32 column indices, int32 or
int64 values.  The reference-shaped classes (gpu.CSRMatrixGPU,
sparse.rb_matrix) wrap these.
"""
from __future__ import annotations

import contextlib
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class DevCSR:
    indptr: object   # torch int64 [n_rows + 1]
    indices: object  # torch int32 [nnz]
    data: object     # torch int32/int64 [nnz]
    shape: tuple

    @property
    def nnz(self) -> int:
        return int(self.indices.numel())

    @property
    def n_rows(self) -> int:
        return int(self.shape[0])


def from_scipy(m, dtype=None) -> DevCSR:
    import scipy.sparse as sp
    m = sp.csr_matrix(m)
    t = _lib.torch()
    dev = _lib.device()
    dt = np.dtype(dtype) if dtype is not None else (np.dtype(np.int64) if m.dtype.itemsize > 4 else np.dtype(np.int32))
    return DevCSR(
        t.from_numpy(np.ascontiguousarray(m.indptr, dtype=np.int64)).to(dev),
        t.from_numpy(np.ascontiguousarray(m.indices, dtype=np.int32)).to(dev),
        t.from_numpy(np.ascontiguousarray(m.data.astype(dt))).to(dev),
        tuple(m.shape))


def to_scipy(c: DevCSR):
    import scipy.sparse as sp
    indptr = c.indptr.cpu().numpy()
    idx_dt = np.int32 if (indptr.size == 0 or indptr[-1] < 2 ** 31) else np.int64
    return sp.csr_matrix((c.data.cpu().numpy(), c.indices.cpu().numpy(), indptr.astype(idx_dt)), shape=c.shape)


def canonicalize(c: DevCSR, stream=None) -> DevCSR:
    """Sorted rows, duplicates combined with the AVOS sum, explicit zeros dropped.

    A matrix that already is canonical is returned as it is (one test kernel
    and one wait instead of a sorted copy)."""
    from .dense import status
    t = _lib.torch()
    n = c.n_rows
    if n == 0:
        return c
    sfx = "i32" if c.data.dtype == t.int32 else "i64"
    sh = _lib.stream_handle(stream)
    st = status()
    st.reset(stream)
    _lib.call(f"rbg_csr_check_canonical_{sfx}", _lib.ptr(c.indptr), _lib.ptr(c.indices), _lib.ptr(c.data), n, st.p,
              sh)
    if st.read(stream).aux[0] == 0:
        return c
    idx = c.indices.clone()
    val = c.data.clone()
    tmp = t.empty(int(_lib.load().rbg_csr_tmp_bytes(n)), dtype=t.uint8, device=idx.device)
    optr = t.empty(n + 1, dtype=t.int64, device=idx.device)
    fn = "rbg_csr_canonicalize_i32" if val.dtype == t.int32 else "rbg_csr_canonicalize_i64"
    _lib.call(fn, _lib.ptr(c.indptr), _lib.ptr(idx), _lib.ptr(val), n, _lib.ptr(tmp), _lib.ptr(optr), None, None, sh)
    nnz = int(optr[-1].item())
    oi = t.empty(nnz, dtype=t.int32, device=idx.device)
    ov = t.empty(nnz, dtype=val.dtype, device=idx.device)
    _lib.call(fn, _lib.ptr(c.indptr), _lib.ptr(idx), _lib.ptr(val), n, _lib.ptr(tmp), _lib.ptr(optr), _lib.ptr(oi),
              _lib.ptr(ov), sh)
    return DevCSR(optr, oi, ov, c.shape)


def validate(c: DevCSR, stream=None) -> bool:
    from .dense import status
    st = status()
    st.reset(stream)
    if c.n_rows:
        _lib.call("rbg_csr_validate", _lib.ptr(c.indptr), _lib.ptr(c.indices), c.n_rows, int(c.shape[1]), st.p,
                  _lib.stream_handle(stream))
    s = st.read(stream)
    return not (s.flags & _lib.F_CAPACITY)


@dataclass
class DagInfo:
    nnz: int
    cand_max: int
    final_max: int
    cycle_vertex: int
    nlevels: int
    events: int
    arena: int = 0      # arena entries the run reserved (sum of candidate counts)

    @property
    def diameter_candidates(self) -> int:
        """_transitive_closure_dag.pyx:347-349: MSB(max |candidate|) when > 1."""
        return int(self.cand_max).bit_length() - 1 if self.cand_max > 1 else 0

    @property
    def diameter_final(self) -> int:
        """gpu/transitive_closure.py:640-646: floor(log2(max |final| > 1))."""
        return int(self.final_max).bit_length() - 1 if self.final_max > 1 else 0


_KINDS_MAX = 4096  # levels whose kinds a plan keeps (deeper graphs run every merge kernel each round)


class DagPlan:
    """Reusable device buffers for closing graphs of one size: the workspace
    and the row arena.  Every level reserves its candidate count in the arena,
    so it needs sum_v (nnz(A[v]) + sum_w nnz(C[w])) entries (DagInfo.arena
    after a run), somewhat more than nnz(C)."""

    def __init__(self, n: int, value_dtype, arena_entries: int, nnz: int = 0):
        t = _lib.torch()
        dev = _lib.device()
        L = _lib.load()
        self.n = n
        self.vdt = value_dtype
        self.nnz = int(nnz)
        self.ws = t.empty(int(L.rbg_dag_workspace_bytes(max(n, 1), self.nnz)), dtype=t.uint8, device=dev)
        self.arena_cols = t.empty(arena_entries, dtype=t.int32, device=dev)
        self.arena_vals = t.empty(arena_entries, dtype=value_dtype, device=dev)
        self.levels = 0     # levels of the last graph closed with this plan (0: unknown)
        self.out_nnz = 0    # closure entries of that graph (0: unknown)
        self.kinds = None   # one byte per level of that graph: which merge kernels it needed
        self.kinds_host = None


def dag_closure(A: DevCSR, width: int | None = None, stream=None, plan: DagPlan | None = None,
                arena_factor: float = 48.0, batch: int = 0):
    """Closure of a DAG held as a canonical device CSR.

    Returns (C: DevCSR, DagInfo).  On a cycle returns (None, info) with
    info.cycle_vertex >= 0.  When a capacity flag comes back the buffers grow
    and the closure runs again.  `batch` is the number of level rounds between
    completion tests (0: the level count of the plan's last graph, queued
    without any test, else 8); with the right count the host waits once per
    closure, for the status and output size together.
    """
    from .dense import status
    t = _lib.torch()
    n = A.n_rows
    vdt = A.data.dtype
    if width is None:
        width = 32 if vdt == t.int32 else 64
    if n == 0:
        return A, DagInfo(0, 0, 0, -1, 0, 0)
    st = status()
    sh = _lib.stream_handle(stream)
    if plan is None or plan.n != n or plan.vdt != vdt or plan.nnz < A.nnz:
        plan = DagPlan(n, vdt, int(max(4096, arena_factor * (A.nnz + n))), A.nnz)
    sfx = "i32" if vdt == t.int32 else "i64"
    dev = A.indices.device
    exact = batch == 0 and plan.levels > 0
    for _attempt in range(8):
        st.reset(stream)
        rounds = -plan.levels if exact else (batch if batch > 0 else 8)
        # the levels' kinds from the last closure with this plan launch only the
        # merge kernels each level needs; a level that needs another one stops
        # the rounds, and the closure runs again with tested rounds
        hinted = exact and plan.kinds is not None and len(plan.kinds) == plan.levels
        kinds = (ctypes.c_uint8 * len(plan.kinds)).from_buffer_copy(plan.kinds) if hinted else None
        _lib.call(f"rbg_dag_closure_{sfx}", _lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data), n,
                  _lib.ptr(plan.ws), plan.ws.numel(), _lib.ptr(plan.arena_cols), _lib.ptr(plan.arena_vals),
                  plan.arena_cols.numel(), rounds, ctypes.cast(kinds, ctypes.c_void_p) if hinted else None, width,
                  st.p, sh)
        # The row pointers are scanned before the status is read, so that one
        # wait gives the flags and the output size together.  When the plan
        # knows its last output size, the entries are gathered into buffers of
        # that size in the same queue (a gather that does not fit writes nothing
        # and says so); otherwise they are gathered after the wait.
        optr = t.empty(n + 1, dtype=t.int64, device=dev)
        cap = plan.out_nnz if exact else 0
        oi = t.empty(cap, dtype=t.int32, device=dev) if cap else None
        ov = t.empty(cap, dtype=vdt, device=dev) if cap else None
        _lib.call(f"rbg_dag_gather_{sfx}", _lib.ptr(plan.ws), n, _lib.ptr(plan.arena_cols),
                  _lib.ptr(plan.arena_vals), _lib.ptr(optr), _lib.ptr(oi) if cap else None,
                  _lib.ptr(ov) if cap else None, cap, st.p, sh)
        if not hinted:
            kcount = min(n, _KINDS_MAX)
            if plan.kinds_host is None or plan.kinds_host.numel() < kcount:
                plan.kinds_host = t.empty(kcount, dtype=t.uint8, pin_memory=True)
            _lib.call("rbg_dag_level_kinds", _lib.ptr(plan.ws), n, _lib.ptr(plan.kinds_host), kcount, sh)
        s = st.read(stream)
        if s.flags & _lib.F_CYCLE:
            return None, DagInfo(0, 0, 0, int(s.aux[0]), int(s.aux[6]), 0)
        nnz = int(s.aux[1])
        if s.flags & _lib.F_CAPACITY and not (s.aux[7] and cap and nnz > cap):
            plan = DagPlan(n, vdt, 2 * plan.arena_cols.numel(), A.nnz)
            exact = False
            continue
        if not s.aux[7]:  # the plan's level count was too small for this graph
            exact = False
            continue
        info = DagInfo(0, int(s.aux[2]), int(s.aux[3]), -1, int(s.aux[6]), int(s.events))
        info.arena = int(s.aux[4])
        plan.levels = info.nlevels
        plan.out_nnz = nnz
        if not hinted:
            plan.kinds = bytes(plan.kinds_host[:info.nlevels].numpy()) if info.nlevels <= _KINDS_MAX else None
        if cap and nnz <= cap:
            oi, ov = oi[:nnz], ov[:nnz]
        else:
            oi = t.empty(nnz, dtype=t.int32, device=dev)
            ov = t.empty(nnz, dtype=vdt, device=dev)
            _lib.call(f"rbg_dag_gather_{sfx}", _lib.ptr(plan.ws), n, _lib.ptr(plan.arena_cols),
                      _lib.ptr(plan.arena_vals), _lib.ptr(optr), _lib.ptr(oi), _lib.ptr(ov), nnz, st.p, sh)
        info.nnz = nnz
        return DevCSR(optr, oi, ov, A.shape), info
    raise MemoryError("DAG closure buffers could not be sized")


def spgemm(A: DevCSR, B: DevCSR, mask_upper: bool = False, width: int | None = None, is_unsigned: bool = False,
           stream=None, arena_factor: float = 4.0):
    """C = A (x) B (AVOS semiring) for canonical device CSR operands.  Returns (C, events)."""
    from .dense import status
    t = _lib.torch()
    m = A.n_rows
    if A.shape[1] != B.shape[0]:
        raise ValueError(f"Incompatible shapes: A is {A.shape}, B is {B.shape}. A.shape[1] must equal B.shape[0].")
    vdt = A.data.dtype
    if B.data.dtype != vdt:
        B = DevCSR(B.indptr, B.indices, B.data.to(vdt), B.shape)
    if width is None:
        width = 32 if vdt == t.int32 else 64
    dev = A.indices.device
    if m == 0 or A.nnz == 0 or B.nnz == 0:
        return DevCSR(t.zeros(m + 1, dtype=t.int64, device=dev), t.empty(0, dtype=t.int32, device=dev),
                      t.empty(0, dtype=vdt, device=dev), (m, B.shape[1])), 0
    # the row bounds sum to the expansion size: sum over A's entries of |B row|
    st = status()
    sh = _lib.stream_handle(stream)
    st.reset(stream)
    _lib.call("rbg_csr_expand_size", _lib.ptr(A.indptr), _lib.ptr(A.indices), m, _lib.ptr(B.indptr), st.p, sh)
    expand = int(st.read(stream).aux[0])
    sfx = "i32" if vdt == t.int32 else "i64"
    ws = t.empty(int(_lib.load().rbg_dag_workspace_bytes(m, 0)), dtype=t.uint8, device=dev)
    cap = max(1024, expand)
    tcap = max(1024, 3 * expand)
    ac = t.empty(cap, dtype=t.int32, device=dev)
    av = t.empty(cap, dtype=vdt, device=dev)
    tc = t.empty(tcap, dtype=t.int32, device=dev)
    tv = t.empty(tcap, dtype=vdt, device=dev)
    st.reset(stream)
    _lib.call(f"rbg_spgemm_rows_{sfx}", _lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data), m,
              _lib.ptr(B.indptr), _lib.ptr(B.indices), _lib.ptr(B.data), int(mask_upper), _lib.ptr(ws),
              _lib.ptr(ac), _lib.ptr(av), cap, _lib.ptr(tc), _lib.ptr(tv), tcap, width, int(is_unsigned), st.p, sh)
    optr = t.empty(m + 1, dtype=t.int64, device=dev)
    _lib.call(f"rbg_dag_gather_{sfx}", _lib.ptr(ws), m, _lib.ptr(ac), _lib.ptr(av), _lib.ptr(optr), None, None, 0,
              st.p, sh)
    s = st.read(stream)
    if s.flags & _lib.F_CAPACITY:
        raise MemoryError("SpGEMM arena too small")  # bounds are exact; not expected
    nnz = int(s.aux[1])
    oi = t.empty(nnz, dtype=t.int32, device=dev)
    ov = t.empty(nnz, dtype=vdt, device=dev)
    _lib.call(f"rbg_dag_gather_{sfx}", _lib.ptr(ws), m, _lib.ptr(ac), _lib.ptr(av), _lib.ptr(optr), _lib.ptr(oi),
              _lib.ptr(ov), nnz, st.p, sh)
    s = st.read(stream)
    return DevCSR(optr, oi, ov, (m, B.shape[1])), int(s.events)


def avos_sum_csr(A: DevCSR, B: DevCSR, min_abs: bool = False, stream=None) -> DevCSR:
    """Element-wise sum of two canonical device CSR matrices of one shape.

    Default: the AVOS sum (nonzero min, RED_ONE absorbing).  min_abs: keep the
    entry of smallest |value| (first operand on ties), the merge rule of the
    reference's squaring closures (gpu/transitive_closure.py:225-257,
    sparse/csgraph/transitive_closure.py:331-344).  Zero results are dropped."""
    t = _lib.torch()
    if A.shape != B.shape:
        raise ValueError(f"Shape mismatch: {A.shape} vs {B.shape}")
    n = A.n_rows
    dev = A.indices.device
    vdt = A.data.dtype
    if n == 0:
        return A
    bdata = B.data.to(vdt)
    tmp = t.empty(int(_lib.load().rbg_csr_tmp_bytes(n)), dtype=t.uint8, device=dev)
    cat_ptr = A.indptr + B.indptr
    cat_idx = t.empty(A.nnz + B.nnz, dtype=t.int32, device=dev)
    cat_val = t.empty(A.nnz + B.nnz, dtype=vdt, device=dev)
    optr = t.empty(n + 1, dtype=t.int64, device=dev)
    fn = "rbg_csr_add_i32" if vdt == t.int32 else "rbg_csr_add_i64"
    sh = _lib.stream_handle(stream)
    args = [_lib.ptr(A.indptr), _lib.ptr(A.indices), _lib.ptr(A.data), _lib.ptr(B.indptr), _lib.ptr(B.indices),
            _lib.ptr(bdata), n, int(min_abs), _lib.ptr(cat_ptr), _lib.ptr(cat_idx), _lib.ptr(cat_val), _lib.ptr(tmp),
            _lib.ptr(optr)]
    _lib.call(fn, *args, None, None, sh)
    nnz = int(optr[-1].item())
    oi = t.empty(nnz, dtype=t.int32, device=dev)
    ov = t.empty(nnz, dtype=vdt, device=dev)
    _lib.call(fn, *args, _lib.ptr(oi), _lib.ptr(ov), sh)
    return DevCSR(optr, oi, ov, A.shape)


def csr_equal(A: DevCSR, B: DevCSR) -> bool:
    """sparse_equal_gpu (gpu/transitive_closure.py:159-184)."""
    t = _lib.torch()
    if A.shape != B.shape or A.nnz != B.nnz:
        return False
    if A.nnz == 0:
        return True
    return bool(t.equal(A.indptr, B.indptr)) and bool(t.equal(A.indices, B.indices)) and \
        bool(t.equal(A.data, B.data.to(A.data.dtype)))


# Largest dense form of a closure that edge_insert_many builds (bytes).
DENSE_INSERT_MAX_BYTES = 4 << 30


def _sfx(dt) -> str:
    return "i32" if dt == _lib.torch().int32 else "i64"


def glue_temp(n: int, device):
    t = _lib.torch()
    return t.empty(int(_lib.load().rbg_glue_temp_bytes(max(n, 1))), dtype=t.uint8, device=device)


def to_dense(C: DevCSR, stream=None):
    """The n x m device matrix of a CSR (absent entries 0): rbg_csr_to_dense."""
    t = _lib.torch()
    n, m = C.shape
    D = t.empty((n, m), dtype=C.data.dtype, device=C.indices.device)
    if n:
        _lib.call(f"rbg_csr_to_dense_{_sfx(C.data.dtype)}", _lib.ptr(C.indptr), _lib.ptr(C.indices), _lib.ptr(C.data),
                  n, m, _lib.ptr(D), _lib.stream_handle(stream))
    return D


def from_dense(D, stream=None) -> DevCSR:
    """Canonical CSR of a device matrix (row-major order, zeros dropped):
    rbg_dense_to_csr, counting pass then filling pass."""
    t = _lib.torch()
    D = D.contiguous()
    n, m = D.shape
    indptr = t.empty(n + 1, dtype=t.int64, device=D.device)
    temp = glue_temp(n, D.device)
    fn = f"rbg_dense_to_csr_{_sfx(D.dtype)}"
    sh = _lib.stream_handle(stream)
    _lib.call(fn, _lib.ptr(D), n, m, _lib.ptr(indptr), None, None, _lib.ptr(temp), temp.numel(), sh)
    nnz = int(indptr[n].item()) if n else 0
    idx = t.empty(nnz, dtype=t.int32, device=D.device)
    val = t.empty(nnz, dtype=D.dtype, device=D.device)
    if nnz:
        _lib.call(fn, _lib.ptr(D), n, m, _lib.ptr(indptr), _lib.ptr(idx), _lib.ptr(val), _lib.ptr(temp), temp.numel(),
                  sh)
    return DevCSR(indptr, idx, val, (n, m))


def column(C: DevCSR, col: int, stream=None) -> DevCSR:
    """Column `col` of a sorted CSR as an n x 1 CSR (rbg_csr_column)."""
    t = _lib.torch()
    n = C.n_rows
    dev = C.indices.device
    optr = t.empty(n + 1, dtype=t.int64, device=dev)
    oi = t.empty(max(n, 1), dtype=t.int32, device=dev)
    ov = t.empty(max(n, 1), dtype=C.data.dtype, device=dev)
    temp = glue_temp(n, dev)
    _lib.call(f"rbg_csr_column_{_sfx(C.data.dtype)}", _lib.ptr(C.indptr), _lib.ptr(C.indices), _lib.ptr(C.data), n,
              int(col), _lib.ptr(optr), _lib.ptr(oi), _lib.ptr(ov), _lib.ptr(temp), temp.numel(),
              _lib.stream_handle(stream))
    return DevCSR(optr, oi, ov, (n, 1))


def edge_insert_many(C: DevCSR, edges, stream=None) -> DevCSR:
    """C with the parent edges (alpha, beta, relationship) inserted one after
    another, each as edge_insert.  While C's dense form fits
    DENSE_INSERT_MAX_BYTES the edges are applied to it by the dense rank-1
    kernels back to back (rbg_edge_compose_*, one status read for all of
    them) and the result goes back to CSR; otherwise edge by edge.  Raises
    ValueError when an edge would close a cycle, OverflowError on an
    overflowing product."""
    from . import dense
    from ._dtypes import elem_type
    t = _lib.torch()
    n = C.n_rows
    edges = [(int(a), int(b), int(r)) for (a, b, r) in edges]
    for a, b, _ in edges:
        if not (0 <= a < n and 0 <= b < n):
            raise ValueError("vertex index out of range")
    if not edges:
        return C
    if n * n * C.data.element_size() > DENSE_INSERT_MAX_BYTES:
        for a, b, r in edges:
            C = edge_insert(C, a, b, r, stream=stream)
        return C
    et = elem_type(np.int32 if C.data.dtype == t.int32 else np.int64)
    with t.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
        D = to_dense(C, stream)
        W, st = dense.edge_compose_many(D, edges, et, stream=stream, inplace=True)
        if st.flags & _lib.F_CYCLE:
            raise ValueError("Relational composition would result in a cycle.")
        if st.events:
            raise OverflowError("Avos product results in an overflow during edge composition")
        return from_dense(W, stream)


def edge_insert(C: DevCSR, alpha: int, beta: int, relationship: int, stream=None) -> DevCSR:
    """Closure of C with the edge alpha -> beta (value ``relationship``) added,
    for a closed canonical device CSR C: the sparse form of the edge
    relational composition (core/redblack.py:74-97,
    relational_composition.c.in:71-137).

    u = (row alpha of C with entry beta set to the relationship) (x) C is the
    new row alpha; every row m then becomes C[m] (+) C[m][alpha] (x) u (the
    row alpha itself included: its diagonal +-1 leaves u unchanged).  Both
    steps run on the GPU as AVOS SpGEMMs and one CSR AVOS sum.  Raises
    ValueError when C[beta][alpha] != 0 (the edge would close a cycle).
    """
    from .dense import status
    t = _lib.torch()
    n = C.n_rows
    if not (0 <= alpha < n and 0 <= beta < n):
        raise ValueError("vertex index out of range")
    dev = C.indices.device
    vdt = C.data.dtype
    sh = _lib.stream_handle(stream)
    # column alpha of C as an n x 1 matrix; C[beta][alpha] != 0 closes a cycle
    X = column(C, alpha, stream)
    xp = X.indptr[beta:beta + 2].cpu()
    if int(xp[1]) > int(xp[0]):
        raise ValueError("Relational composition would result in a cycle.")
    # the composed row's seed: row alpha of C with entry beta = relationship
    a_len = int((C.indptr[alpha + 1] - C.indptr[alpha]).item())
    cols = t.empty(a_len + 1, dtype=t.int32, device=dev)
    vals = t.empty(a_len + 1, dtype=vdt, device=dev)
    st = status()
    st.reset(stream)
    _lib.call(f"rbg_csr_row_set_{_sfx(vdt)}", _lib.ptr(C.indptr), _lib.ptr(C.indices), _lib.ptr(C.data), alpha, beta,
              int(relationship), _lib.ptr(cols), _lib.ptr(vals), st.p, sh)
    k = int(st.read(stream).aux[0])
    row = DevCSR(t.tensor([0, k], dtype=t.int64, device=dev), cols[:k], vals[:k], (1, n))
    u, _ = spgemm(row, C, stream=stream)
    nx = int(X.indptr[n].item())
    X = DevCSR(X.indptr, X.indices[:nx], X.data[:nx], (n, 1))
    R1, _ = spgemm(X, u, stream=stream)
    return avos_sum_csr(C, R1, stream=stream)
