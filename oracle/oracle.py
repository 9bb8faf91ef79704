"""CPU oracle for the AVOS closure path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this module, and only as the checker (or the timed CPU
comparator).  The HIP library never calls it and never falls back to it.

What each function restates (reference file:line) and what pins it
independently of itself:

=====================  ====================================================  =========================================
function               restates                                              pinned by (tests/test_oracle.py)
=====================  ====================================================  =========================================
avos_sum / product     rbg_math.c.in:29-121 (C core), rbm_math.h:13-110      17x17 / 5x5 tables of tests/avos/test_math.py:23-68;
                                                                             64-bit edge + overflow cases (:87-111);
                                                                             compiled reference rbm_math.h (oracle/_ref);
                                                                             closed form ((x-1)<<MSB(y))+y
warshall (naive FW)    warshall.c.in:50-107                                  5x5/7x7 closures and diameters of
                                                                             tests/avos/test_apsp.py:23-106; Ahnentafel
                                                                             BFS on valid pedigrees; brute-force
                                                                             first-discovery path enumeration for the
                                                                             trajectory diameter; cycle example of
                                                                             tests/avos/test_cycle_detection.py:40-52
cy_floyd_warshall      _shortest_path.pyx:340-389                            same 5x5/7x7 answers; equality with the
                                                                             DAG closure on the seeded random DAGs of
                                                                             tests/sparse/test_dag_sparse_closure.py
matmul                 redblack.c.src:554-700 ('ij,jk' loop)                 tests/test_redblack.py:19-74 known answers;
                                                                             brute-force mat_avos restatement
edge/vertex compose    relational_composition.c.in:51-267                    tests/avos/test_relational_composition.py
                                                                             known answers + exact cycle message
spgemm pass1/pass2     rbm_math.h:116-222                                    compiled reference rbm_math.h (oracle/_ref);
                                                                             dense triple loop
topological_sort       _topological_sort.pyx:30-163                          order validity on random DAGs; the
                                                                             reference's cycle example
dag_closure            _transitive_closure_dag.pyx:168-351                   tests/sparse/test_dag_sparse_closure.py and
                                                                             tests/reference/test_transitive_closure_dag.py
                                                                             known answers; FW equality on the seeded
                                                                             DAGs; Ahnentafel BFS
canonical_ordering     reference/ordering.py:13-53, components.py:4-52       tests/reference/test_triangularization.py
                       (_ordering.pyx:17-47, :160-180)                        golden permutation and matrix; component
                                                                             grouping of tests/sparse/test_ordering.py;
                                                                             brute-force key check on random closures
=====================  ====================================================  =========================================

Parity unpinned (no reference vector or closed form reaches it): the exact
text of the *last* cycle-detection message on large cyclic inputs, overflow
events on narrow dtypes beyond the reference's two overflow cases, and the
Cython DAG closure's candidate-max diameter on graphs with pedigree collapse
(checked only against brute force on small graphs).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_NAMES = {
    np.dtype(np.int8): "byte", np.dtype(np.uint8): "ubyte",
    np.dtype(np.int16): "short", np.dtype(np.uint16): "ushort",
    np.dtype(np.int32): "int", np.dtype(np.uint32): "uint",
    np.dtype(np.int64): "long", np.dtype(np.uint64): "ulong",
}
_CT = {
    "byte": ctypes.c_int8, "ubyte": ctypes.c_uint8, "short": ctypes.c_int16,
    "ushort": ctypes.c_uint16, "int": ctypes.c_int32, "uint": ctypes.c_uint32,
    "long": ctypes.c_int64, "ulong": ctypes.c_uint64,
}
_UCT = {
    "byte": ctypes.c_uint8, "ubyte": ctypes.c_uint8, "short": ctypes.c_uint16,
    "ushort": ctypes.c_uint16, "int": ctypes.c_uint32, "uint": ctypes.c_uint32,
    "long": ctypes.c_uint64, "ulong": ctypes.c_uint64,
}

OE_WIDTH, OE_RED, OE_HANG = 1, 2, 4


class FwReport(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "n_cycle", "first_k", "first_i", "first_a", "first_b",
        "last_k", "last_i", "last_a", "last_b",
        "n_overflow", "ov_k", "ov_i", "ov_j", "ov_a", "ov_b", "ov_kind",
        "last_is_overflow", "hang", "diameter")]


class DagResult(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64),
                ("indptr", ctypes.POINTER(ctypes.c_int64)),
                ("indices", ctypes.POINTER(ctypes.c_int32)),
                ("data", ctypes.POINTER(ctypes.c_int64)),
                ("cand_max", ctypes.c_int64), ("final_max", ctypes.c_int64),
                ("cycle_vertex", ctypes.c_int64), ("events", ctypes.c_int64)]


def build() -> None:
    """Compile liborc.so (and oracle/_ref when the reference tree exists)."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE, "liborc.so"])
    if os.path.isdir("/root/reference/redblackgraph"):
        subprocess.check_call(["make", "-s", "-C", _HERE, "ref"])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        for name, ct in _CT.items():
            getattr(L, f"orc_{name}_avos_sum").restype = ct
            getattr(L, f"orc_{name}_avos_sum").argtypes = [ct, ct]
            getattr(L, f"orc_{name}_avos_product").restype = _UCT[name]
            getattr(L, f"orc_{name}_avos_product").argtypes = [ct, ct, ctypes.POINTER(ctypes.c_int)]
            getattr(L, f"orc_{name}_warshall").argtypes = [P, ctypes.c_int64, ctypes.POINTER(FwReport)]
            getattr(L, f"orc_{name}_matmul").argtypes = [P, P, P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
            getattr(L, f"orc_{name}_matmul").restype = ctypes.c_int
        for name in ("int", "long"):
            ct = _CT[name]
            f = getattr(L, f"orc_{name}_edge_compose")
            f.argtypes = [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ct, P]
            f.restype = ctypes.c_int
            f = getattr(L, f"orc_{name}_vertex_compose")
            f.argtypes = [P, P, P, ct, ctypes.c_int64, P, ctypes.POINTER(ctypes.c_int)]
            f.restype = ctypes.c_int64
        L.orc_cy_floyd_warshall.argtypes = [P, ctypes.c_int64, ctypes.c_int, ctypes.POINTER(FwReport)]
        L.orc_cy_floyd_warshall.restype = ctypes.c_int
        L.orc_cy_avos_product.argtypes = [ctypes.c_int32, ctypes.c_int32]
        L.orc_cy_avos_product.restype = ctypes.c_int32
        L.orc_spgemm_pass1.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, P, P, P]
        L.orc_spgemm_pass1.restype = ctypes.c_int64
        L.orc_spgemm_pass2.argtypes = [ctypes.c_int32, ctypes.c_int32, P, P, P, P, P, P, P, P, P]
        L.orc_spgemm_pass2.restype = ctypes.c_int
        L.orc_topological_sort.argtypes = [ctypes.c_int32, P, P, P]
        L.orc_topological_sort.restype = ctypes.c_int64
        L.orc_dag_closure.argtypes = [ctypes.c_int32, P, P, P, ctypes.c_int]
        L.orc_dag_closure.restype = ctypes.POINTER(DagResult)
        L.orc_dag_free.argtypes = [ctypes.POINTER(DagResult)]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own rbm_math.h compiled by oracle/Makefile, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "librbm_ref.so")
        if not os.path.exists(path):
            return None
        L = ctypes.PyDLL(path)
        P = ctypes.c_void_p
        L.ref_avos_product_ll.argtypes = [ctypes.c_longlong, ctypes.c_longlong]
        L.ref_avos_product_ll.restype = ctypes.c_ulonglong
        L.ref_avos_sum_ll.argtypes = [ctypes.c_longlong, ctypes.c_longlong]
        L.ref_avos_sum_ll.restype = ctypes.c_longlong
        L.ref_avos_product_i.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ref_avos_product_i.restype = ctypes.c_uint
        L.ref_avos_sum_i.argtypes = [ctypes.c_int, ctypes.c_int]
        L.ref_avos_sum_i.restype = ctypes.c_int
        L.ref_matmat_pass1.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P]
        L.ref_matmat_pass2.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, P, P, P]
        _REF = L
    return _REF


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------------------
# scalar arithmetic
# ---------------------------------------------------------------------------
def avos_sum(a: int, b: int, dtype=np.int64) -> int:
    name = _NAMES[np.dtype(dtype)]
    ct = _CT[name]
    return int(getattr(lib(), f"orc_{name}_avos_sum")(ct(int(a)).value, ct(int(b)).value))


def avos_product(a: int, b: int, dtype=np.int64):
    """Returns (result as the element type, events)."""
    name = _NAMES[np.dtype(dtype)]
    ct = _CT[name]
    ev = ctypes.c_int(0)
    r = getattr(lib(), f"orc_{name}_avos_product")(ct(int(a)).value, ct(int(b)).value, ctypes.byref(ev))
    # reinterpret the unsigned result as the element type (C conversion)
    return int(np.array([r], dtype=np.uint64).astype(np.dtype(dtype))[0]), ev.value


def avos_product_unsigned(a: int, b: int) -> tuple:
    """The scalar entry point's view: 64-bit unsigned result (c_avos_product_impl)."""
    ev = ctypes.c_int(0)
    r = lib().orc_long_avos_product(int(a), int(b), ctypes.byref(ev))
    return int(r), ev.value


def cy_avos_product(a: int, b: int) -> int:
    return int(lib().orc_cy_avos_product(int(a), int(b)))


# ---------------------------------------------------------------------------
# dense closure / product
# ---------------------------------------------------------------------------
@dataclass
class FwResult:
    W: np.ndarray
    diameter: int
    report: dict


def warshall(A, dtype=None) -> FwResult:
    """Naive FW exactly as the C core runs it; W and report (errors not raised)."""
    A = np.array(A, dtype=dtype if dtype is not None else np.asarray(A).dtype, copy=True)
    A = np.ascontiguousarray(A)
    name = _NAMES[A.dtype]
    rep = FwReport()
    getattr(lib(), f"orc_{name}_warshall")(_ptr(A), A.shape[0], ctypes.byref(rep))
    d = {f: getattr(rep, f) for f, _ in FwReport._fields_}
    return FwResult(A, int(rep.diameter), d)


def cy_floyd_warshall(A, upper: bool = False) -> FwResult:
    A = np.ascontiguousarray(np.array(A, dtype=np.int32, copy=True))
    rep = FwReport()
    rc = lib().orc_cy_floyd_warshall(_ptr(A), A.shape[0], int(upper), ctypes.byref(rep))
    d = {f: getattr(rep, f) for f, _ in FwReport._fields_}
    d["raised"] = bool(rc)
    return FwResult(A, int(rep.diameter), d)


def matmul(A, B):
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B, dtype=A.dtype)
    M, K = A.shape
    K2, N = B.shape
    assert K == K2
    C = np.zeros((M, N), dtype=A.dtype)
    ev = getattr(lib(), f"orc_{_NAMES[A.dtype]}_matmul")(_ptr(A), _ptr(B), _ptr(C), M, K, N)
    return C, ev


def edge_compose(R, alpha, beta, rel):
    R = np.ascontiguousarray(R)
    out = np.empty_like(R)
    ev = getattr(lib(), f"orc_{_NAMES[R.dtype]}_edge_compose")(_ptr(R), R.shape[0], alpha, beta, int(rel), _ptr(out))
    return out, ev


def vertex_compose(u, R, v, color):
    R = np.ascontiguousarray(R)
    n = R.shape[0]
    u = np.ascontiguousarray(np.asarray(u, dtype=R.dtype).reshape(n))
    v = np.ascontiguousarray(np.asarray(v, dtype=R.dtype).reshape(n))
    out = np.zeros((n + 1, n + 1), dtype=R.dtype)
    ev = ctypes.c_int(0)
    rc = getattr(lib(), f"orc_{_NAMES[R.dtype]}_vertex_compose")(_ptr(u), _ptr(R), _ptr(v), int(color), n, _ptr(out), ctypes.byref(ev))
    return out, int(rc), ev.value


# ---------------------------------------------------------------------------
# sparse
# ---------------------------------------------------------------------------
def spgemm(A, B):
    """Returns (indptr, indices, data) in the reference's pass-2 column order."""
    A = A.tocsr(); B = B.tocsr()
    Ap = np.ascontiguousarray(A.indptr, dtype=np.int32); Aj = np.ascontiguousarray(A.indices, dtype=np.int32)
    Bp = np.ascontiguousarray(B.indptr, dtype=np.int32); Bj = np.ascontiguousarray(B.indices, dtype=np.int32)
    Ax = np.ascontiguousarray(A.data, dtype=np.int64); Bx = np.ascontiguousarray(B.data, dtype=np.int64)
    M, N = A.shape[0], B.shape[1]
    Cp = np.zeros(M + 1, dtype=np.int64)
    nnz = lib().orc_spgemm_pass1(M, N, _ptr(Ap), _ptr(Aj), _ptr(Bp), _ptr(Bj), _ptr(Cp))
    Cj = np.zeros(max(nnz, 1), dtype=np.int32); Cx = np.zeros(max(nnz, 1), dtype=np.int64)
    ev = lib().orc_spgemm_pass2(M, N, _ptr(Ap), _ptr(Aj), _ptr(Ax), _ptr(Bp), _ptr(Bj), _ptr(Bx), _ptr(Cp), _ptr(Cj), _ptr(Cx))
    nnz2 = int(Cp[-1])
    return Cp, Cj[:nnz2], Cx[:nnz2], ev


def topological_sort(A):
    A = A.tocsr()
    n = A.shape[0]
    ip = np.ascontiguousarray(A.indptr, dtype=np.int32); ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    order = np.zeros(max(n, 1), dtype=np.int32)
    bad = lib().orc_topological_sort(n, _ptr(ip), _ptr(ix), _ptr(order))
    return order[:n], int(bad)


@dataclass
class DagOut:
    indptr: np.ndarray
    indices: np.ndarray
    data: np.ndarray
    cand_max: int
    final_max: int
    cycle_vertex: int
    events: int

    @property
    def diameter_cython(self) -> int:
        # _transitive_closure_dag.pyx:347-349: MSB(max |candidate|) if > 1
        return int(self.cand_max).bit_length() - 1 if self.cand_max > 1 else 0

    @property
    def diameter_final(self) -> int:
        # gpu/transitive_closure.py:640-646: floor(log2(max |final| > 1))
        return int(self.final_max).bit_length() - 1 if self.final_max > 1 else 0

    def toarray(self, n):
        import scipy.sparse as sp
        return sp.csr_matrix((self.data, self.indices, self.indptr), shape=(n, n)).toarray()


def dag_closure(A, width: int = 64) -> DagOut:
    A = A.tocsr()
    n = A.shape[0]
    ip = np.ascontiguousarray(A.indptr, dtype=np.int32); ix = np.ascontiguousarray(A.indices, dtype=np.int32)
    dx = np.ascontiguousarray(A.data, dtype=np.int64)
    L = lib()
    r = L.orc_dag_closure(n, _ptr(ip), _ptr(ix), _ptr(dx), width)
    try:
        res = r.contents
        if res.cycle_vertex >= 0:
            return DagOut(None, None, None, 0, 0, int(res.cycle_vertex), int(res.events))
        nnz = res.nnz
        indptr = np.ctypeslib.as_array(res.indptr, shape=(n + 1,)).copy()
        indices = np.ctypeslib.as_array(res.indices, shape=(max(nnz, 1),))[:nnz].copy()
        data = np.ctypeslib.as_array(res.data, shape=(max(nnz, 1),))[:nnz].copy()
        return DagOut(indptr, indices, data, int(res.cand_max), int(res.final_max), -1, int(res.events))
    finally:
        L.orc_dag_free(r)


# ---------------------------------------------------------------------------
# pure-Python restatements for tiny inputs
# ---------------------------------------------------------------------------
def py_msb(x: int) -> int:
    """reference/rbg_math.py:1-11"""
    b = 0
    while x > 1:
        x >>= 1
        b += 1
    return b


def py_avos_sum(x: int, y: int) -> int:
    """reference/rbg_math.py:13-26"""
    if x == 0:
        return y
    if y == 0:
        return x
    return x if x < y else y


def py_avos_product(x: int, y: int) -> int:
    """reference/rbg_math.py:28-93"""
    if x < -1 or y < -1:
        raise ValueError("invalid operand")
    if x == 0 or y == 0:
        return 0
    if x == -1 and y == -1:
        return -1
    if x == 1 and y == 1:
        return 1
    if (x, y) in ((-1, 1), (1, -1)):
        return 0
    if x == -1:
        x = 1
    if y == -1:
        return 0 if x & 1 else x
    if y == 1:
        return x if x & 1 else 0
    bp = py_msb(y)
    return (y & (2 ** bp - 1)) | (x << bp)


def py_mat_avos(A, B):
    """reference/mat_avos.py: element (i,k) = avos-sum over j of A[i][j] (x) B[j][k]"""
    A = [[int(v) for v in row] for row in A]
    B = [[int(v) for v in row] for row in B]
    n, m, p = len(A), len(B), len(B[0])
    out = [[0] * p for _ in range(n)]
    for i in range(n):
        for k in range(p):
            acc = 0
            for j in range(m):
                acc = py_avos_sum(acc, py_avos_product(A[i][j], B[j][k]))
            out[i][k] = acc
    return out


def canonical_ordering(A):
    """reference/ordering.py:13-53 (avos_canonical_ordering, _get_permutation)
    with reference/components.py:4-52 (find_components) and
    reference/permutation.py:3-15 (permute), in numpy.  Returns
    (B, permutation, {component id: size})."""
    A = np.asarray(A, dtype=np.int64)
    n = A.shape[0]
    adj = (A != 0) | (A != 0).T
    comp = np.full(n, -1, dtype=np.int64)
    sizes = {}
    cid = 0
    for s0 in range(n):  # components in order of their smallest vertex
        if comp[s0] >= 0:
            continue
        stack = [s0]
        comp[s0] = cid
        cnt = 0
        while stack:
            v = stack.pop()
            cnt += 1
            for w in np.nonzero(adj[v])[0]:
                if comp[w] < 0:
                    comp[w] = cid
                    stack.append(int(w))
        sizes[cid] = cnt
        cid += 1
    msb = np.vectorize(py_msb, otypes=[np.int64])
    anc = np.where(A != 0, msb(A), 0).sum(axis=1) if n else np.zeros(0, np.int64)
    maxrel = np.maximum(A.max(axis=0), 0) if n else np.zeros(0, np.int64)
    diag = np.diagonal(A)
    basis = sorted(range(n), key=lambda x: (-sizes[comp[x]], comp[x], -anc[x], maxrel[x], -diag[x], x))
    p = np.array(basis, dtype=np.int64)
    B = A[np.ix_(p, p)] if n else A.copy()
    return B, p, sizes


def canonical_ordering_csr(A):
    """sparse/csgraph/_ordering.pyx:50-157 (_get_permutation_sparse and
    _compute_ordering_metrics_sparse) in vectorised numpy/scipy, for CSR
    closures too large for canonical_ordering.  Components come from
    scipy.sparse.csgraph.connected_components (weak), renumbered by each
    component's smallest vertex as in the reference.  Returns
    (permutation, {component id: size})."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    A = sp.csr_matrix(A)
    n = A.shape[0]
    _, lab = connected_components(A, directed=True, connection="weak")
    first = np.full(lab.max() + 1 if n else 0, n, dtype=np.int64)
    np.minimum.at(first, lab, np.arange(n))
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    comp = rank[lab]
    sizes_arr = np.bincount(comp, minlength=order.size)
    data = A.data.astype(np.int64)
    nz = data != 0
    msb = np.zeros(data.size, dtype=np.int64)
    big = data > 1
    msb[big] = np.floor(np.log2(data[big].astype(np.float64))).astype(np.int64)
    fix = big & (np.left_shift(np.int64(1), msb, where=big, out=np.zeros_like(msb)) > data)
    msb[fix] -= 1  # exact floor(log2) where the float rounding overshoots
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    anc = np.bincount(rows, weights=np.where(nz, msb, 0), minlength=n).astype(np.int64)
    maxrel = np.zeros(n, dtype=np.int64)
    np.maximum.at(maxrel, A.indices[nz], data[nz])
    diag = np.zeros(n, dtype=np.int64)
    on = A.indices == rows
    diag[rows[on]] = data[on]
    # np.lexsort: the last key is the primary one
    perm = np.lexsort((np.arange(n), -diag, maxrel, -anc, comp, -sizes_arr[comp]))
    return perm, {int(k): int(v) for k, v in enumerate(sizes_arr)}
