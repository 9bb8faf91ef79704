"""GPU tests of component finding and the component-wise closure: the
reference's own cases (tests/sparse/test_component_closure.py,
tests/sparse/test_components.py) plus oracle comparisons on shuffled
multi-pedigree graphs."""
import numpy as np
import pytest
import scipy.sparse as sp

from redblackgraph_amd import _lib, synth
from redblackgraph_amd.sparse import csgraph as cg

pytestmark = pytest.mark.gpu


def _union_find_ids(A):
    """Component ids numbered by smallest vertex (plain union-find)."""
    A = sp.coo_matrix(A)
    n = A.shape[0]
    parent = list(range(n))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x
    for r, c in zip(A.row, A.col):
        a, b = find(int(r)), find(int(c))
        if a != b:
            parent[max(a, b)] = min(a, b)
    roots = np.array([find(v) for v in range(n)])
    return np.unique(roots, return_inverse=True)[1]


def test_reference_two_components():
    A = sp.csr_matrix(np.array([[1, 2, 0, 0], [0, 1, 0, 0], [0, 0, 1, 4], [0, 0, 0, 1]], dtype=np.int32))
    res = cg.component_wise_closure(A)
    assert sp.isspmatrix(res.W)
    W = res.W.toarray()
    assert not W[:2, 2:].any() and not W[2:, :2].any()
    np.testing.assert_array_equal(W, A.toarray())
    for m in ("FW", "D"):
        assert sp.isspmatrix(cg.component_wise_closure(A, method=m).W)


def test_reference_matches_full_closure(orc):
    A = np.array([[1, 2, 0, 0, 0, 0], [0, 1, 4, 0, 0, 0], [0, 0, 1, 0, 0, 0],
                  [0, 0, 0, 1, 2, 0], [0, 0, 0, 0, 1, 0], [0, 0, 0, 0, 0, 1]], dtype=np.int32)
    res = cg.component_wise_closure(sp.csr_matrix(A))
    ref = orc.warshall(A)
    np.testing.assert_array_equal(res.W.toarray(), ref.W)
    assert res.diameter == ref.diameter
    assert res.W.toarray()[0, 2] == 8


def test_reference_singletons_empty_single():
    n = 100
    res = cg.component_wise_closure(sp.csr_matrix(np.eye(n, dtype=np.int32)))
    np.testing.assert_array_equal(res.W.toarray(), np.eye(n, dtype=np.int32))
    res = cg.component_wise_closure(sp.csr_matrix((0, 0), dtype=np.int32))
    assert res.W.shape == (0, 0) and res.diameter == 0
    res = cg.component_wise_closure(sp.csr_matrix([[1]], dtype=np.int32))
    assert res.W.shape == (1, 1) and res.W.toarray()[0, 0] == 1


def test_reference_block_diagonal_random(orc):
    """tests/sparse/test_component_closure.py:173-206: 20 random upper
    triangular blocks of 10 (legacy RandomState seed 42)."""
    np.random.seed(42)
    n, k = 200, 20
    size = n // k
    A = np.zeros((n, n), dtype=np.int32)
    for i in range(k):
        block = np.triu(np.random.randint(0, 4, size=(size, size)))
        np.fill_diagonal(block, 1)
        A[i * size:(i + 1) * size, i * size:(i + 1) * size] = block
    res = cg.component_wise_closure(sp.csr_matrix(A))
    ref = orc.warshall(A)
    np.testing.assert_array_equal(res.W.toarray(), ref.W)
    assert res.diameter == ref.diameter


def test_find_components_numbering():
    rng = np.random.default_rng(3)
    n = 3000
    rows = rng.integers(0, n, 2200)
    cols = rng.integers(0, n, 2200)
    A = sp.csr_matrix((np.full(2200, 2, dtype=np.int32), (rows, cols)), shape=(n, n))
    q = {}
    ids = cg.find_components_sparse(A, q)
    exp = _union_find_ids(A)
    np.testing.assert_array_equal(ids, exp)
    assert ids.dtype == np.uint32
    assert sum(q.values()) == n and len(q) == exp.max() + 1
    np.testing.assert_array_equal(cg.find_components(A.toarray()), exp)
    verts = cg.get_component_vertices(ids)
    assert all(np.all(np.diff(v) > 0) for v in verts.values())


def test_shuffled_pedigrees_vs_oracle(orc):
    """60 pedigrees of 16..400 vertices, vertices shuffled together: the
    component-wise closure equals each pedigree's oracle closure, placed back
    at the shuffled indices; the diameter is the largest one."""
    sizes = synth.batched_sizes(60, 16, 400, 9)
    mats = [synth.dense_pedigree(synth.component_spec(int(s)), 500 + i) for i, s in enumerate(sizes)]
    n = int(sizes.sum())
    perm = np.random.default_rng(1).permutation(n)
    rows, cols, vals = [], [], []
    exp_r, exp_c, exp_v = [], [], []
    off, dmax = 0, 0
    for A in mats:
        m = A.shape[0]
        r, c = np.nonzero(A)
        rows.append(perm[off + r]); cols.append(perm[off + c]); vals.append(A[r, c])
        ref = orc.warshall(A)
        dmax = max(dmax, ref.diameter)
        r, c = np.nonzero(ref.W)
        exp_r.append(perm[off + r]); exp_c.append(perm[off + c]); exp_v.append(ref.W[r, c])
        off += m
    G = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    E = sp.csr_matrix((np.concatenate(exp_v), (np.concatenate(exp_r), np.concatenate(exp_c))), shape=(n, n))
    res = cg.component_wise_closure(G)
    E.sort_indices()
    assert (res.W != E).nnz == 0
    np.testing.assert_array_equal(res.W.indptr, E.indptr)
    np.testing.assert_array_equal(res.W.indices, E.indices)
    assert res.diameter == dmax
    auto = cg.transitive_closure_adaptive(G)
    assert (auto.W != E).nnz == 0


def test_cycle_in_one_component(orc):
    """A cycle in the second component raises CycleError at its first
    detection, vertex numbered within that component."""
    a = synth.dense_pedigree(synth.component_spec(40), 1)
    b = synth.dense_pedigree(synth.component_spec(30), 2)
    r, c = np.nonzero(np.tril(b, -1))
    b[c[0], r[0]] = 2
    n = 70
    A = np.zeros((n, n), dtype=np.int32)
    A[:40, :40] = a
    A[40:, 40:] = b
    rep = orc.warshall(b).report
    with pytest.raises(cg.CycleError) as e:
        cg.component_wise_closure(sp.csr_matrix(A))
    assert e.value.vertex == rep["first_i"]


# ---------------------------------------------------------------------------
# device helpers of the component-wise closure and the CSR paths
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_dense_csr_round_trip(dt):
    """rbg_csr_to_dense / rbg_dense_to_csr against scipy, rows of every length
    class (empty, < 64, > 64 entries)."""
    from redblackgraph_amd import sparse_dev as sd
    rng = np.random.default_rng(4)
    n, m = 700, 530
    D = np.where(rng.random((n, m)) < 0.05, rng.integers(-1, 1 << 20, (n, m)), 0).astype(dt)
    D[10] = 0
    D[11, :200] = rng.integers(2, 9, 200)
    C = sd.from_dense(_lib.to_device(D))
    ref = sp.csr_matrix(D)
    ref.sort_indices()
    np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
    np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
    np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
    np.testing.assert_array_equal(sd.to_dense(C).cpu().numpy(), D)


@pytest.mark.parametrize("seed", [0, 1])
def test_components_layout_scatter_merge(seed):
    """rbg_components_layout against numpy (ids by smallest vertex, local
    order, members, sizes, starts); scatter then merge returns the graph."""
    from redblackgraph_amd import sparse_dev as sd
    t = _lib.torch()
    rng = np.random.default_rng(seed)
    blocks = [synth.dense_pedigree(synth.PedigreeSpec(int(k), 4, 0.8, 2), seed * 7 + i)
              for i, k in enumerate(rng.integers(1, 90, 30))]
    A = sp.block_diag(blocks).tocsr()
    p = rng.permutation(A.shape[0])
    A = sp.csr_matrix(A.toarray()[np.ix_(p, p)].astype(np.int32))
    n = A.shape[0]
    dA = sd.from_scipy(A, np.int32)
    from redblackgraph_amd.sparse.csgraph._components import _device_labels_dev
    labels = _device_labels_dev(dA)
    dev = labels.device
    comp, local, members, sizes = (t.empty(n, dtype=t.int32, device=dev) for _ in range(4))
    starts = t.empty(n + 1, dtype=t.int64, device=dev)
    ncomp = t.empty(1, dtype=t.int64, device=dev)
    temp = sd.glue_temp(n, dev)
    _lib.call("rbg_components_layout", _lib.ptr(labels), n, _lib.ptr(comp), _lib.ptr(local), _lib.ptr(members),
              _lib.ptr(sizes), _lib.ptr(starts), _lib.ptr(ncomp), None, _lib.ptr(temp), temp.numel(), _lib.stream_handle())
    nc, lab = sp.csgraph.connected_components(A, directed=True, connection="weak")
    first = np.full(nc, n)
    np.minimum.at(first, lab, np.arange(n))
    rank = np.empty(nc, dtype=np.int64)
    rank[np.argsort(first)] = np.arange(nc)
    ids = rank[lab]
    assert int(ncomp.item()) == nc
    np.testing.assert_array_equal(comp.cpu().numpy(), ids)
    mem = np.argsort(ids, kind="stable")
    np.testing.assert_array_equal(members.cpu().numpy(), mem)
    sz = np.bincount(ids, minlength=nc)
    np.testing.assert_array_equal(sizes.cpu().numpy()[:nc], sz)
    st = np.concatenate([[0], np.cumsum(sz)])
    np.testing.assert_array_equal(starts.cpu().numpy()[:nc + 1], st)
    loc = np.empty(n, dtype=np.int64)
    loc[mem] = np.arange(n) - st[ids[mem]]
    np.testing.assert_array_equal(local.cpu().numpy(), loc)
    # scatter into padded matrices, then merge back: the graph itself
    npad = ((sz + 127) // 128 * 128).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(npad.astype(np.int64) ** 2)[:-1]])
    pool = t.zeros(int((npad.astype(np.int64) ** 2).sum()), dtype=t.int32, device=dev)
    offd, npd = _lib.to_device(off.astype(np.int64)), _lib.to_device(npad)
    sh = _lib.stream_handle()
    _lib.call("rbg_components_scatter", _lib.ptr(dA.indptr), _lib.ptr(dA.indices), _lib.ptr(dA.data), n,
              _lib.ptr(comp), _lib.ptr(local), _lib.ptr(offd), _lib.ptr(npd), _lib.ptr(pool), sh)
    indptr = t.empty(n + 1, dtype=t.int64, device=dev)
    args = [_lib.ptr(pool), _lib.ptr(offd), _lib.ptr(npd), _lib.ptr(comp), _lib.ptr(local), _lib.ptr(members),
            _lib.ptr(starts), _lib.ptr(sizes), n, _lib.ptr(indptr)]
    _lib.call("rbg_components_merge", *args, None, None, _lib.ptr(temp), temp.numel(), sh)
    nnz = int(indptr[n].item())
    oi = t.empty(nnz, dtype=t.int32, device=dev)
    ov = t.empty(nnz, dtype=t.int32, device=dev)
    _lib.call("rbg_components_merge", *args, _lib.ptr(oi), _lib.ptr(ov), _lib.ptr(temp), temp.numel(), sh)
    ref = A.copy()
    ref.sort_indices()
    np.testing.assert_array_equal(indptr.cpu().numpy(), ref.indptr)
    np.testing.assert_array_equal(oi.cpu().numpy(), ref.indices)
    np.testing.assert_array_equal(ov.cpu().numpy(), ref.data)
