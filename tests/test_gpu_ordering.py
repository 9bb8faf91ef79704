"""Canonical ordering and permutation on the GPU (sparse/csgraph/_ordering.pyx,
_permutation.pyx) against the oracle's restatement of reference/ordering.py
and the reference's golden triangularization."""
import numpy as np
import pytest
import scipy.sparse as sp

from redblackgraph_amd import synth

pytestmark = pytest.mark.gpu


def _cg():
    from redblackgraph_amd.sparse import csgraph
    return csgraph


def test_golden_triangularization(golden):
    g = golden["canonical_ordering_7"]
    for A in (np.array(g["A_star"], np.int32), sp.csr_matrix(np.array(g["A_star"], np.int32))):
        o = _cg().avos_canonical_ordering(A)
        np.testing.assert_array_equal(o.label_permutation, g["permutation"])
        B = o.A.toarray() if sp.isspmatrix(o.A) else o.A
        np.testing.assert_array_equal(B, g["canonical"])
        assert o.components == {0: 4, 1: 3}
    g2 = golden["ordering_two_components"]
    o = _cg().avos_canonical_ordering(sp.csr_matrix(np.array(g2["A"], np.int32)))
    assert len(o.components) == g2["n_components"] and sum(o.components.values()) == g2["total"]


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
@pytest.mark.parametrize("seed", [0, 1])
def test_ordering_matches_oracle_dense_and_sparse(orc, dtype, seed):
    """Pedigree closures of several components (the reference's own
    precondition: A transitively closed): dense and CSR input give the
    oracle's permutation and permuted matrix."""
    rng = np.random.default_rng(seed)
    blocks = []
    for k in range(4):
        n = int(rng.integers(20, 120))
        A = synth.dense_pedigree(synth.PedigreeSpec(n, 5, 0.8, 2), seed * 10 + k)
        blocks.append(np.asarray(orc.warshall(A).W))
    C = sp.block_diag(blocks).toarray().astype(dtype)
    perm_idx = rng.permutation(C.shape[0])  # interleave the components' vertices
    C = C[np.ix_(perm_idx, perm_idx)]
    B_ref, p_ref, q_ref = orc.canonical_ordering(C)
    od = _cg().avos_canonical_ordering(C)
    np.testing.assert_array_equal(od.label_permutation, p_ref)
    np.testing.assert_array_equal(od.A, B_ref)
    assert od.components == q_ref
    os_ = _cg().avos_canonical_ordering(sp.csr_matrix(C))
    np.testing.assert_array_equal(os_.label_permutation, p_ref)
    np.testing.assert_array_equal(os_.A.toarray(), B_ref)
    assert os_.A.has_sorted_indices


def test_ordering_reference_random_closure(orc):
    """tests/sparse/test_ordering.py:44-75: a random upper triangular matrix
    closed by Floyd-Warshall; dense and sparse orders agree and match the
    oracle."""
    np.random.seed(42)
    n = 30
    A = np.triu(np.random.randint(0, 4, size=(n, n)), k=1)
    np.fill_diagonal(A, np.random.choice([-1, 1], size=n))
    W = np.asarray(orc.cy_floyd_warshall(A.astype(np.int32)).W).astype(np.int32)
    _, p_ref, _ = orc.canonical_ordering(W)
    pd = _cg().avos_canonical_ordering(W).label_permutation
    ps = _cg().avos_canonical_ordering(sp.csr_matrix(W)).label_permutation
    np.testing.assert_array_equal(pd, p_ref)
    np.testing.assert_array_equal(ps, p_ref)


@pytest.mark.parametrize("upper", [False, True])
def test_permute_dense_and_sparse(upper):
    rng = np.random.default_rng(7)
    n = 300
    A = (rng.random((n, n)) < 0.05) * rng.integers(1, 50, (n, n))
    A = A.astype(np.int64)
    p = rng.permutation(n).astype(np.int32)
    ref = A[np.ix_(p, p)]
    if upper:
        ref = np.triu(ref)
    np.testing.assert_array_equal(_cg().permute(A, p, upper), ref)
    S = _cg().permute_sparse(sp.csr_matrix(A), p, upper)
    assert S.has_sorted_indices
    np.testing.assert_array_equal(S.toarray(), ref)


def test_ordering_sparse_pedigree_closure(orc):
    """A 1,500-vertex pedigree closed on the GPU, ordered as CSR: the
    permutation matches the oracle's (computed on the dense closure) and the
    result is upper triangular."""
    from redblackgraph_amd import sparse_dev as sd
    A = synth.csr_pedigree(synth.PedigreeSpec(1500, 8, 0.7, 1), 4)
    C, _ = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    Cs = sd.to_scipy(C)
    _, p_ref, q_ref = orc.canonical_ordering(Cs.toarray())
    o = _cg().avos_canonical_ordering(Cs)
    np.testing.assert_array_equal(o.label_permutation, p_ref)
    assert o.components == q_ref
    assert sp.triu(o.A).nnz == o.A.nnz


def test_canonical_order_device_matches_host_entry(orc):
    """The device-only entry (DevCSR in, DevCSR out) gives the same
    permutation, matrix and component sizes as avos_canonical_ordering."""
    from redblackgraph_amd import sparse_dev as sd
    from redblackgraph_amd.sparse.csgraph._ordering import canonical_order_device
    A = synth.csr_pedigree(synth.PedigreeSpec(4000, 9, 0.7, 1), 8)
    C, _ = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    perm, B, counts = canonical_order_device(C)
    Cs = sd.to_scipy(C)
    ref_perm, ref_q = orc.canonical_ordering_csr(Cs)
    np.testing.assert_array_equal(perm.cpu().numpy(), ref_perm)
    assert {i: int(c) for i, c in enumerate(counts.cpu().numpy())} == ref_q
    Bs = sd.to_scipy(B)
    R = Cs[ref_perm][:, ref_perm]
    R.sort_indices()
    np.testing.assert_array_equal(Bs.indptr, R.indptr)
    np.testing.assert_array_equal(Bs.indices, R.indices)
    np.testing.assert_array_equal(Bs.data, R.data)


@pytest.mark.parametrize("dtype", [np.int32, np.int64])
def test_permute_sparse_long_rows(dtype):
    """Rows of more than 256 entries take the staging buffer and the
    segmented sort; short rows among them take the one-pass ranking.  Both
    against scipy's A[p][:, p] with sorted indices, entry for entry."""
    rng = np.random.default_rng(21)
    n = 6000
    A = sp.random(n, n, density=0.002, format="lil", random_state=7,
                  data_rvs=lambda k: rng.integers(2, 1 << 20, k))
    for r, length in ((11, 257), (900, 1024), (4000, 5000), (5999, 300)):
        cols = rng.choice(n, size=length, replace=False)
        A[r, cols] = rng.integers(2, 1 << 20, length)
    A = sp.csr_matrix(A, dtype=dtype)
    A.sum_duplicates()
    lens = np.diff(A.indptr)
    assert lens.max() >= 5000 and (lens > 256).sum() >= 4 and (lens <= 256).sum() > n - 10
    p = rng.permutation(n).astype(np.int32)
    B = _cg().permute_sparse(A, p)
    ref = A[p][:, p].tocsr()
    ref.sort_indices()
    assert B.has_sorted_indices
    np.testing.assert_array_equal(B.indptr, ref.indptr)
    np.testing.assert_array_equal(B.indices, ref.indices)
    np.testing.assert_array_equal(B.data, ref.data.astype(dtype))
    # the triangular filter on the same input
    Bu = _cg().permute_sparse(A, p, assume_upper_triangular=True)
    refu = sp.triu(ref).tocsr()
    refu.sort_indices()
    np.testing.assert_array_equal(Bu.indptr, refu.indptr)
    np.testing.assert_array_equal(Bu.indices, refu.indices)
    np.testing.assert_array_equal(Bu.data, refu.data.astype(dtype))


@pytest.mark.parametrize("L,need_gib", [((1 << 15) + 1, 100), ((1 << 16) + 1, 130)])
def test_permute_csr_past_2_31_entries(L, need_gib):
    """The CSR boundary past 2^31 entries: 65,536 rows of L entries, rows
    unsorted, through rbg_permute_csr_i32 with a random permutation.
    L = 32,769: 2,147,549,184 entries, row pointers and output offsets past
    2^31, long rows sorted by rocPRIM's segmented sort (32-bit sizes).
    L = 65,537: 4,295,032,832 entries, past 2^32, so the long rows take the
    library's own 64-bit row sort; each row holds one column twice (k = 0 and
    k = n), and the two keep their order.  Checked on the device, row block
    by row block: every row has its L entries, columns non-decreasing, and
    each entry is the input entry it names (row p[i], column p[j], value a
    function of its place in the input row).  About 70 and 110 GB of device
    memory."""
    from redblackgraph_amd import _lib
    from redblackgraph_amd.sparse.csgraph._ordering import _permute_csr_device
    t = _lib.torch()
    dev = _lib.device()
    free, _ = t.cuda.mem_get_info(dev)
    if free < need_gib * 2 ** 30:
        pytest.skip(f"needs about {need_gib} GiB of free device memory")
    n, b = 1 << 16, 40503  # b odd: k -> k b mod n is one-to-one on [0, n)
    binv = pow(b, -1, n)
    nnz = n * L
    assert nnz > 2 ** 31
    kk = t.arange(L, device=dev, dtype=t.int64)
    indptr = t.arange(n + 1, device=dev, dtype=t.int64) * L
    indices = t.empty(nnz, dtype=t.int32, device=dev)
    data = t.empty(nnz, dtype=t.int32, device=dev)
    R = 1024
    for r0 in range(0, n, R):  # row r: columns (r + k b) mod n, values 2 + (7 r + 3 k) mod 1000
        r = t.arange(r0, r0 + R, device=dev, dtype=t.int64)[:, None]
        indices[r0 * L:(r0 + R) * L] = ((r + kk[None, :] * b) % n).to(t.int32).reshape(-1)
        data[r0 * L:(r0 + R) * L] = (2 + (7 * r + 3 * kk[None, :]) % 1000).to(t.int32).reshape(-1)
    del kk
    g = t.Generator(device=dev)
    g.manual_seed(5)
    p = t.randperm(n, device=dev, generator=g).to(t.int32)
    optr, oidx, oval = _permute_csr_device(indptr, indices, data, n, nnz, np.dtype(np.int32), p)
    del indices, data
    t.cuda.synchronize()
    assert t.equal(optr, indptr)
    p64 = p.to(t.int64)
    for i0 in range(0, n, R):
        e0, e1 = i0 * L, (i0 + R) * L
        cols = oidx[e0:e1].to(t.int64).reshape(R, L)
        step = cols[:, 1:] - cols[:, :-1]
        assert bool((step >= 0).all())
        pi = p64[i0:i0 + R][:, None]
        k = ((p64[cols] - pi) % n) * binv % n          # place in input row p[i], modulo n
        if L > n:                                      # the second of two equal columns is k + n
            assert int((step == 0).sum()) == R
            k[:, 1:] += (step == 0).to(t.int64) * n
        else:
            assert bool((step > 0).all())
        assert bool((k < L).all())
        want = (2 + (7 * pi + 3 * k) % 1000).to(t.int32)
        assert t.equal(oval[e0:e1].reshape(R, L), want)
