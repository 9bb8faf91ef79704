"""GPU parity of the csgraph entry points that sit on the closure path:
topological_sort / topological_ordering / is_upper_triangular
(_topological_sort.pyx:30-265), floyd_warshall / shortest_path
(_shortest_path.pyx:36-389) and the csgraph relational compositions
(_relational_composition.pyx:37-113).  The cases follow the reference's
tests/sparse/test_topological_sort.py; closures are compared with the
oracle's restatement of the Cython loop."""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import redblackgraph_amd as rb
from redblackgraph_amd import synth
from redblackgraph_amd.sparse import csgraph as cg
from redblackgraph_amd.sparse.csgraph import CycleError

pytestmark = pytest.mark.gpu


def _edges_forward(A, p):
    """Every edge i -> j (i != j) of A goes forward under the order p."""
    A = sp.csr_matrix(A)
    pos = np.empty(A.shape[0], dtype=np.int64)
    pos[np.asarray(p)] = np.arange(A.shape[0])
    r, c = A.nonzero()
    off = r != c
    return bool(np.all(pos[r[off]] < pos[c[off]]))


def _is_perm(p, n):
    return len(p) == n and np.array_equal(np.sort(p), np.arange(n))


# ---------------------------------------------------------------------------
# topological sort
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("M", [
    [[1, 2, 0], [0, 1, 4], [0, 0, 1]],            # a chain 0 -> 1 -> 2, already upper triangular
    [[1, 0, 0], [2, 1, 0], [4, 2, 1]],            # lower triangular: must be reversed
    [[1, 2, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4], [0, 0, 0, 1]],  # two components
    [[1, 2, 0, 0, 0, 0, 0], [0, 1, 4, 0, 0, 0, 0], [0, 0, 1, 0, 0, 0, 0], [0, 0, 0, 1, 0, 0, 0],
     [0, 0, 0, 0, 1, 0, 0], [0, 0, 0, 0, 0, 1, 2], [0, 0, 0, 0, 0, 0, 1]],  # mixed component sizes
])
def test_topological_sort_small(M):
    A = sp.csr_matrix(np.array(M, dtype=np.int32))
    p = cg.topological_sort(A)
    assert p.dtype == np.int32 and _is_perm(p, A.shape[0])
    assert _edges_forward(A, p)
    B = cg.permute_sparse(A, p, assume_upper_triangular=True)
    assert cg.is_upper_triangular(B)
    assert B.nnz == A.nnz  # nothing fell below the diagonal


def test_topological_sort_inputs():
    M = np.array([[1, 2, 4], [0, 1, 2], [0, 0, 1]], dtype=np.int32)
    for A in (M, sp.csr_matrix(M), sp.csr_matrix(M).tocsc()):
        p = cg.topological_sort(A)
        assert _is_perm(p, 3) and _edges_forward(M, p)
    assert len(cg.topological_sort(sp.csr_matrix((0, 0), dtype=np.int32))) == 0
    np.testing.assert_array_equal(cg.topological_sort(sp.csr_matrix([[1]], dtype=np.int32)), [0])
    p = cg.topological_sort(sp.csr_matrix(np.eye(100, dtype=np.int32)))  # singletons; self-loops are no edges
    assert _is_perm(p, 100)


def test_topological_sort_random_permuted_upper():
    """An upper triangular DAG renumbered at random: the sort recovers an
    order that makes it upper triangular again."""
    rng = np.random.RandomState(42)
    n = 50
    D = np.triu(rng.randint(0, 5, size=(n, n)), k=0)
    np.fill_diagonal(D, rng.choice([-1, 1], size=n))
    A = sp.csr_matrix(D.astype(np.int32))
    Ap = cg.permute_sparse(A, rng.permutation(n).astype(np.int32))
    p = cg.topological_sort(Ap)
    assert _edges_forward(Ap, p)
    assert cg.is_upper_triangular(cg.permute_sparse(Ap, p, assume_upper_triangular=True))


@pytest.mark.parametrize("M,cycle", [
    ([[1, 2, 0], [0, 1, 4], [8, 0, 1]], {0, 1, 2}),
    ([[1, 2], [4, 1]], {0, 1}),
])
def test_topological_sort_cycles(M, cycle):
    with pytest.raises(CycleError) as e:
        cg.topological_sort(sp.csr_matrix(np.array(M, dtype=np.int32)))
    assert e.value.vertex in cycle


def test_topological_sort_cycle_in_pedigree():
    """A back edge in a 5000-vertex pedigree: the reported vertex lies on a cycle."""
    A = synth.csr_pedigree(synth.PedigreeSpec(5000, 12, 0.8, 2), 3, dtype=np.int32)
    anc = sp.csgraph.breadth_first_order(A, 3000, directed=True, return_predecessors=False)
    far = int(anc[-1])  # an ancestor of 3000 (edges run child -> parent)
    assert far != 3000
    A = A.tolil()
    A[far, 3000] = 2  # ... that now has 3000 as a parent: a cycle
    A = A.tocsr()
    with pytest.raises(CycleError) as e:
        cg.topological_sort(A)
    v = e.value.vertex
    back = sp.csgraph.breadth_first_order(A, v, directed=True, return_predecessors=False)
    # v reaches itself: some successor of v reaches v
    succ = [int(w) for w in A[v].indices if w != v]
    assert any(v in set(sp.csgraph.breadth_first_order(A, w, directed=True, return_predecessors=False).tolist())
               for w in succ), (v, len(back))


def test_topological_sort_large_chain():
    """100,000 vertices: a chain plus 5,000 random forward edges (the
    reference's large-graph case)."""
    rng = np.random.RandomState(42)
    n = 100_000
    row = list(range(n - 1))
    col = list(range(1, n))
    for _ in range(5000):
        i = rng.randint(0, n - 1)
        row.append(i)
        col.append(rng.randint(i + 1, n))
    A = sp.csr_matrix((np.ones(len(row), dtype=np.int32) * 2, (row, col)), shape=(n, n))
    t0 = time.perf_counter()
    p = cg.topological_sort(A)
    assert time.perf_counter() - t0 < 5.0
    assert _edges_forward(A, p)


def test_topological_sort_pedigree_levels():
    """A 200,000-vertex pedigree (lower triangular): every child before its
    parents, and the levels are longest paths to a founder."""
    from redblackgraph_amd.sparse.csgraph._topological_sort import topological_levels
    A = synth.csr_pedigree(synth.PedigreeSpec(200_000, 15, 0.7, 1), 5, dtype=np.int64)
    p, lev = topological_levels(A)
    assert _edges_forward(A, p)
    # longest path to a vertex without parents, by dynamic programming in vertex order
    ref = np.zeros(A.shape[0], dtype=np.int64)
    ip, ix = A.indptr, A.indices
    for v in range(A.shape[0]):
        ps = ix[ip[v]:ip[v + 1]]
        ps = ps[ps != v]
        if ps.size:
            ref[v] = ref[ps].max() + 1
    np.testing.assert_array_equal(lev, ref)


def test_topological_ordering():
    A = sp.csr_matrix(np.array([[1, 0, 0], [2, 1, 0], [4, 2, 1]], dtype=np.int32))
    r = cg.topological_ordering(A)
    from redblackgraph_amd.types import Ordering
    assert isinstance(r, Ordering)
    assert sp.isspmatrix(r.A) and cg.is_upper_triangular(r.A)
    assert r.A.nnz <= A.nnz and r.components == {0: 3}
    D = np.array([[1, 2, 4], [0, 1, 2], [0, 0, 1]], dtype=np.int32)
    r = cg.topological_ordering(D)
    assert not sp.isspmatrix(r.A) and cg.is_upper_triangular(r.A)


@pytest.mark.parametrize("seed", range(4))
def test_is_upper_triangular(seed):
    rng = np.random.default_rng(seed)
    n = 300
    D = np.where(rng.random((n, n)) < 0.02, rng.integers(2, 9, (n, n)), 0).astype(np.int32)
    for M in (np.triu(D), np.triu(D, 1), D, np.tril(D)):
        below = bool(np.any(np.tril(M, -1)))
        on = bool(np.any(np.diagonal(M)))
        assert cg.is_upper_triangular(M) == (not below)
        assert cg.is_upper_triangular(sp.csr_matrix(M), strict=True) == (not below and not on)
    assert cg.is_upper_triangular(sp.csr_matrix((0, 0), dtype=np.int32))


# ---------------------------------------------------------------------------
# floyd_warshall / shortest_path
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n,reverse", [(300, False), (300, True), (700, True)])
def test_floyd_warshall_vs_cython_oracle(orc, n, reverse):
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 30 + n)
    if reverse:
        A = synth.reverse_vertices(A)
    ref = orc.cy_floyd_warshall(A, upper=reverse)
    W, d = cg.floyd_warshall(A, assume_upper_triangular=reverse)
    np.testing.assert_array_equal(W, ref.W)
    assert d == ref.diameter
    W2, d2 = cg.floyd_warshall(sp.csr_matrix(A))
    np.testing.assert_array_equal(W2, ref.W)
    assert d2 == ref.diameter
    W3, P, d3 = cg.floyd_warshall(A, return_predecessors=True)
    np.testing.assert_array_equal(W3, ref.W)
    assert P.shape == A.shape and P.dtype == np.int32


def test_floyd_warshall_cycle_first_detection(orc):
    A = synth.dense_pedigree(synth.PedigreeSpec(260, 10, 0.8, 3), 5)
    rng = np.random.default_rng(5)
    for _ in range(6):
        i, j = sorted(rng.integers(0, 260, 2))
        if i != j:
            A[i, j] = 2 if A[j, j] == -1 else 3
    ref = orc.warshall(A).report
    with pytest.raises(CycleError) as e:
        cg.floyd_warshall(A)
    i, k = ref["first_i"], ref["first_k"]
    assert e.value.vertex == i
    assert f"Vertex {i} has a path to itself. A({i},{k})={ref['first_a']}, A({k},{i})={ref['first_b']}" in str(e.value)


def test_floyd_warshall_overwrite_and_unweighted(golden):
    A = np.array(golden["apsp_5"]["A"], dtype=np.int32)
    W, d = cg.floyd_warshall(A.copy())
    np.testing.assert_array_equal(W, golden["apsp_5"]["W"])
    assert d == golden["apsp_5"]["diameter"]
    B = A.copy()
    W, _ = cg.floyd_warshall(B, overwrite=True)
    np.testing.assert_array_equal(B, golden["apsp_5"]["W"])  # written into the caller's int32 array
    C = A.copy()
    cg.floyd_warshall(C)
    np.testing.assert_array_equal(C, A)  # not without overwrite
    W, d = cg.floyd_warshall(A, unweighted=True)
    assert np.all(W == 1) and d == 0


def test_shortest_path_methods(golden, orc):
    A = synth.dense_pedigree(synth.component_spec(200, 0.8, 3), 8)
    ref = orc.cy_floyd_warshall(A)
    for method in ("auto", "FW", "D", "BF", "J"):
        W, d = cg.shortest_path(A, method=method)
        np.testing.assert_array_equal(W, ref.W)
        assert d == ref.diameter
    rows, d = cg.shortest_path(A, method="D", indices=[3, -1])
    np.testing.assert_array_equal(rows, ref.W[[3, 199]])
    row, d = cg.shortest_path(A, method="D", indices=5)
    np.testing.assert_array_equal(row, ref.W[5])
    with pytest.raises(ValueError):
        cg.shortest_path(A, method="FW", indices=[0])
    with pytest.raises(ValueError):
        cg.shortest_path(A, method="XX")


# ---------------------------------------------------------------------------
# relational composition (Cython versions: no checks)
# ---------------------------------------------------------------------------
def test_csgraph_compositions_known(golden):
    for key in ("vertex_compose_1", "vertex_compose_2", "vertex_compose_14"):
        g = golden[key]
        out = cg.vertex_relational_composition(np.array(g["u"]), np.array(g["A"], dtype=np.int32),
                                               np.array(g["v"]), g["color"])
        np.testing.assert_array_equal(out, g["out"])
        assert type(out) is np.ndarray and out.dtype == np.int32
    for key in ("edge_compose_simple", "edge_compose_15"):
        g = golden[key]
        R = np.array(g["R"], dtype=np.int32)
        keep = R.copy()
        out = cg.edge_relational_composition(R, g["alpha"], g["beta"], g["rel"])
        np.testing.assert_array_equal(out, g["out"])
        np.testing.assert_array_equal(R, keep)


def test_csgraph_compositions_vs_oracle(orc):
    A = synth.dense_pedigree(synth.PedigreeSpec(180, 8, 0.8, 3), 12)
    R = orc.warshall(A).W
    rng = np.random.default_rng(4)
    for _ in range(3):
        a, b = int(rng.integers(90, 180)), int(rng.integers(0, 90))
        rel = 2 if A[b, b] == -1 else 3
        exp, _ = orc.edge_compose(R, a, b, rel)
        np.testing.assert_array_equal(cg.edge_relational_composition(R, a, b, rel), exp)
    u = np.zeros(180, dtype=np.int32)
    v = np.zeros(180, dtype=np.int32)
    u[[3, 40]] = [2, 3]
    v[[100, 150]] = [2, 3]
    exp, _, _ = orc.vertex_compose(u, R, v, -1)
    np.testing.assert_array_equal(cg.vertex_relational_composition(u, R, v, -1), exp)
    # an edge that closes a cycle is composed, not raised (the core version raises)
    i, j = [int(x) for x in np.argwhere(np.tril(R, -1) > 0)[0]]
    exp, _ = orc.edge_compose(R, j, i, 2)
    np.testing.assert_array_equal(cg.edge_relational_composition(R, j, i, 2), exp)
