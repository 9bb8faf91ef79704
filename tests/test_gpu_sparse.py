"""GPU parity of the sparse path: DAG closure (csgraph and gpu entry points),
AVOS SpGEMM, squaring closure, CSRMatrixGPU plumbing -- against the oracle,
the reference's known answers and the Ahnentafel rule.  All compute goes
through librbg_hip.so."""
import numpy as np
import pytest
import scipy.sparse as sp

import redblackgraph_amd as rb
from redblackgraph_amd import synth
from redblackgraph_amd import sparse_dev as sd
from tests.helpers import ahnentafel_closure, random_valid_pedigree

pytestmark = pytest.mark.gpu


def _canon(m):
    m = sp.csr_matrix(m)
    m.eliminate_zeros()
    m.sort_indices()
    return m


def _assert_csr_equal(a, b):
    a, b = _canon(a), _canon(b)
    assert a.shape == b.shape
    np.testing.assert_array_equal(a.indptr, b.indptr)
    np.testing.assert_array_equal(a.indices, b.indices)
    np.testing.assert_array_equal(a.data.astype(np.int64), b.data.astype(np.int64))


# ---------------------------------------------------------------------------
# DAG closure
# ---------------------------------------------------------------------------
def test_dag_known_answers(golden):
    from redblackgraph_amd.sparse.csgraph import transitive_closure_dag_sparse, CycleError
    g = golden
    W = transitive_closure_dag_sparse(sp.csr_matrix(np.array(g["dag_chain"]["A"], np.int32))).W.toarray()
    assert W[0, 2] == g["dag_chain"]["W02"]
    W = transitive_closure_dag_sparse(sp.csr_matrix(np.array(g["dag_known_4"]["A"], np.int32))).W.toarray()
    np.testing.assert_array_equal(W, g["dag_known_4"]["W"])
    W = transitive_closure_dag_sparse(sp.csr_matrix(np.array(g["dag_diamond"]["A"], np.int32))).W.toarray()
    assert W[0, 3] == g["dag_diamond"]["W03"]
    res = transitive_closure_dag_sparse(sp.csr_matrix(np.array(g["dag_parity"]["A"], np.int32)))
    np.testing.assert_array_equal(np.diag(res.W.toarray()), g["dag_parity"]["diag"])
    assert sp.isspmatrix(res.W)
    with pytest.raises(CycleError):
        transitive_closure_dag_sparse(sp.csr_matrix(np.array(g["dag_cycle"]["A"], np.int32)))


def test_dag_edge_cases():
    from redblackgraph_amd.sparse.csgraph import transitive_closure_dag_sparse
    r = transitive_closure_dag_sparse(sp.csr_matrix((0, 0), dtype=np.int32))
    assert r.W.shape == (0, 0)
    assert transitive_closure_dag_sparse(sp.csr_matrix([[1]], dtype=np.int32)).W.toarray()[0, 0] == 1
    I = np.eye(10, dtype=np.int32)
    np.testing.assert_array_equal(transitive_closure_dag_sparse(sp.csr_matrix(I)).W.toarray(), I)
    A = np.eye(5, dtype=np.int32)
    for i in range(4):
        A[i, i + 1] = 2
    W = transitive_closure_dag_sparse(sp.csr_matrix(A)).W.toarray()
    assert all(W[i, j] != 0 for i in range(5) for j in range(i + 1, 5))


@pytest.mark.parametrize("seed,n", [(42, 20), (123, 15), (456, 15)])
def test_dag_matches_fw_on_reference_seeds(orc, seed, n):
    from redblackgraph_amd.sparse.csgraph import transitive_closure_dag_sparse, transitive_closure_adaptive
    A = synth.random_upper_dag(n, seed)
    fw = orc.cy_floyd_warshall(A)
    res = transitive_closure_dag_sparse(sp.csr_matrix(A))
    np.testing.assert_array_equal(res.W.toarray(), fw.W)
    ref = orc.dag_closure(sp.csr_matrix(A), 32)
    assert res.diameter == ref.diameter_cython
    np.testing.assert_array_equal(transitive_closure_adaptive(sp.csr_matrix(A), sparse_only=True).W.toarray(), fw.W)


@pytest.mark.parametrize("seed", range(4))
def test_dag_ahnentafel_and_diameters(orc, seed):
    rng = np.random.default_rng(700 + seed)
    A = random_valid_pedigree(300, rng)
    Acsr = sp.csr_matrix(A)
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(Acsr, np.int64)))
    np.testing.assert_array_equal(sd.to_scipy(C).toarray(), ahnentafel_closure(A))
    ref = orc.dag_closure(Acsr, 64)
    assert info.diameter_candidates == ref.diameter_cython
    assert info.diameter_final == ref.diameter_final


def test_canonicalize_fast_path_and_sorting(orc):
    """A canonical matrix comes back as it is; one with unsorted rows,
    duplicate columns (combined with the AVOS sum) and explicit zeros is
    sorted into a canonical copy."""
    import torch
    from redblackgraph_amd import _lib
    A = synth.csr_pedigree(synth.PedigreeSpec(5000, 6, 0.7, 1), 2)
    Ad = sd.from_scipy(A, np.int64)
    assert sd.canonicalize(Ad) is Ad
    rng = np.random.default_rng(11)
    n = 400
    rows, cols, vals = [], [], []
    for i in range(n):
        k = int(rng.integers(0, 9))
        for c in rng.integers(0, n, k):
            rows.append(i)
            cols.append(int(c))
            vals.append(int(rng.choice([0, -1, 1, 2, 3, 5, 9])))
    indptr = np.zeros(n + 1, np.int64)
    np.add.at(indptr, np.asarray(rows, np.int64) + 1, 1)
    indptr = np.cumsum(indptr)
    order = np.argsort(np.asarray(rows), kind="stable")
    idx = np.asarray(cols, np.int32)[order]
    val = np.asarray(vals, np.int64)[order]
    dev = _lib.device()
    M = sd.DevCSR(torch.from_numpy(indptr).to(dev), torch.from_numpy(idx).to(dev), torch.from_numpy(val).to(dev),
                  (n, n))
    Mc = sd.canonicalize(M)
    assert Mc is not M
    exp_ptr, exp_idx, exp_val = [0], [], []
    for i in range(n):
        acc = {}
        for c, v in zip(idx[indptr[i]:indptr[i + 1]], val[indptr[i]:indptr[i + 1]]):
            acc[int(c)] = orc.avos_sum(acc[int(c)], int(v)) if int(c) in acc else int(v)
        for c in sorted(acc):
            if acc[c] != 0:
                exp_idx.append(c)
                exp_val.append(acc[c])
        exp_ptr.append(len(exp_idx))
    np.testing.assert_array_equal(Mc.indptr.cpu().numpy(), exp_ptr)
    np.testing.assert_array_equal(Mc.indices.cpu().numpy(), exp_idx)
    np.testing.assert_array_equal(Mc.data.cpu().numpy(), exp_val)
    assert sd.canonicalize(Mc) is Mc


@pytest.mark.parametrize("hint", [1, 3, 40])
def test_dag_plan_level_hint(orc, hint):
    """A plan's level count queues the rounds without completion tests; a
    count too small for the graph falls back to tested rounds, one too large
    only adds idle rounds.  Results match the oracle either way."""
    import torch
    A = synth.csr_pedigree(synth.PedigreeSpec(20_000, 12, 0.7, 1), 5)
    Ad = sd.canonicalize(sd.from_scipy(A, np.int64))
    ref = orc.dag_closure(A, 64)
    plan = sd.DagPlan(A.shape[0], torch.int64, 64 * (A.nnz + A.shape[0]), A.nnz)
    plan.levels = hint
    for _ in range(2):
        C, info = sd.dag_closure(Ad, plan=plan)
        np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
        np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
        np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
        assert plan.levels == info.nlevels
    # a cycle under a hint still reports a vertex on it
    Cy = sp.csr_matrix(np.array([[1, 2, 0], [0, 1, 3], [2, 0, 1]], np.int64))
    cplan = sd.DagPlan(3, torch.int64, 4096, Cy.nnz)
    cplan.levels = hint
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(Cy, np.int64)), plan=cplan)
    assert C is None and info.cycle_vertex in (0, 1, 2)


@pytest.mark.parametrize("delta", [-1, 0, 5000])
def test_dag_plan_output_size_hint(orc, delta):
    """With a known output size the entries are gathered before the status is
    read.  Buffers one entry short are not written and the gather is redone
    after the wait; larger ones are trimmed.  A second graph of the same size
    with fewer entries reuses the plan.  Results match the oracle throughout."""
    import torch
    spec = synth.PedigreeSpec(20_000, 12, 0.7, 1)
    graphs = [synth.csr_pedigree(spec, seed) for seed in (5, 6)]
    refs = [orc.dag_closure(A, 64) for A in graphs]
    big = int(np.argmax([r.indices.size for r in refs]))
    order = [big, 1 - big]
    A0 = graphs[order[0]]
    plan = sd.DagPlan(A0.shape[0], torch.int64, 64 * (A0.nnz + A0.shape[0]), max(A.nnz for A in graphs))
    for k, g in enumerate(order):
        Ad = sd.canonicalize(sd.from_scipy(graphs[g], np.int64))
        if k == 0:
            C, info = sd.dag_closure(Ad, plan=plan)  # sizes the plan
            assert plan.out_nnz == info.nnz
            plan.out_nnz = info.nnz + delta
        C, info = sd.dag_closure(Ad, plan=plan)
        ref = refs[g]
        assert info.nnz == ref.indices.size == C.indices.numel() == C.data.numel()
        np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
        np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
        np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
        assert plan.out_nnz == info.nnz
    # a cycle under both hints still reports a vertex on it
    Cy = sp.csr_matrix(np.array([[1, 2, 0], [0, 1, 3], [2, 0, 1]], np.int64))
    cplan = sd.DagPlan(3, torch.int64, 4096, Cy.nnz)
    cplan.levels, cplan.out_nnz = 2, 16
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(Cy, np.int64)), plan=cplan)
    assert C is None and info.cycle_vertex in (0, 1, 2)


def test_dag_plan_level_kinds(orc):
    """A plan records which merge kernels each level needed and the next
    closure of the graph launches only those.  Hints that leave out a kernel a
    level needs (either merge kernel swapped, long rows left out) stop the
    rounds and the closure reruns with tested rounds; extra long-row launches
    are harmless.  Results match the oracle every time."""
    import torch
    star = np.eye(30, dtype=np.int64)
    star[0, 1:22] = 2  # vertex 0 has 22 direct entries: a long row in level 1
    cases = [synth.csr_pedigree(synth.PedigreeSpec(100_000, 15, 0.7, 1), 3), sp.csr_matrix(star)]
    for A in cases:
        ref = orc.dag_closure(A, 64)
        Ad = sd.canonicalize(sd.from_scipy(A, np.int64))
        plan = sd.DagPlan(A.shape[0], torch.int64, 64 * (A.nnz + A.shape[0]), A.nnz)
        C, info = sd.dag_closure(Ad, plan=plan)
        true = plan.kinds
        assert true is not None and len(true) == info.nlevels == plan.levels
        if A.shape[0] == 30:
            assert true == bytes([1, 3])
        else:
            assert any(k & 1 for k in true) and any(not k & 1 for k in true)
        hints = [true, bytes(k ^ 1 for k in true), bytes(k & 1 for k in true), bytes(k | 2 for k in true)]
        for hint in hints:
            plan.kinds = hint
            C, info = sd.dag_closure(Ad, plan=plan)
            np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
            np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
            np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
            needed = all((h & 1) == (k & 1) and not (k & ~h & 2) for h, k in zip(hint, true))
            assert plan.kinds == (hint if needed else true)


def test_dag_pedigree_100k_vs_oracle(orc):
    spec = synth.PedigreeSpec(100_000, 15, 0.7, 1)
    A = synth.csr_pedigree(spec, 3)
    ref = orc.dag_closure(A, 64)
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
    np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
    np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
    assert (info.cand_max, info.final_max) == (ref.cand_max, ref.final_max)


def test_dag_1m_workload_vs_oracle(orc):
    """The 1M-vertex sparse workload (15 layers, p=0.7, int64 values): the
    whole closure bit-exact against the oracle."""
    A = synth.csr_pedigree(synth.SPARSE_1M, 0)
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    ref = orc.dag_closure(A, 64)
    assert info.nnz == ref.indptr[-1]
    np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
    np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
    np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
    assert info.diameter_candidates == ref.diameter_cython
    # sortedness and diagonal preserved
    d = C.data.cpu().numpy()
    assert np.all(np.abs(d) >= 1)


@pytest.mark.parametrize("vdt,width", [(np.int64, 64), (np.int32, 32)])
def test_dag_long_rows_all_merge_paths(orc, vdt, width):
    """Rows of every merge path: vertex 0 points at 2999 vertices (more than
    the workgroup's 2048-entry LDS buffers: merged in its arena slot), vertex 1
    at 1200 (workgroup in LDS), vertex 2 at 300 (one wave in LDS), and vertex 3
    reaches vertices 0..2's rows through successors 1 and 2."""
    n = 3000
    rng = np.random.default_rng(4)
    A = np.zeros((n, n), dtype=np.int64)
    np.fill_diagonal(A, rng.choice([-1, 1], n))
    A[0, 1:] = rng.integers(2, 4, n - 1)
    A[1, 4:1204] = rng.integers(2, 4, 1200)
    A[2, 1000:1300] = rng.integers(2, 4, 300)
    A[3, [1, 2]] = [2, 3]
    A[3, 2500:2600] = 2
    Acsr = sp.csr_matrix(A)
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(Acsr, vdt)))
    ref = orc.dag_closure(Acsr, width)
    assert ref.events == 0
    np.testing.assert_array_equal(sd.to_scipy(C).toarray(), ref.toarray(n))
    assert info.diameter_candidates == ref.diameter_cython
    assert (info.cand_max, info.final_max) == (ref.cand_max, ref.final_max)


def test_dag_wide_fanout(orc):
    """Vertices with many successors (a layered DAG, not a pedigree)."""
    rng = np.random.default_rng(9)
    layers, width, fan = 8, 50, 12
    n = layers * width
    A = np.zeros((n, n), dtype=np.int64)
    np.fill_diagonal(A, rng.choice([-1, 1], n))
    for L in range(layers - 1):
        for i in range(L * width, (L + 1) * width):
            for j in rng.choice(np.arange((L + 1) * width, (L + 2) * width), size=fan, replace=False):
                A[i, j] = int(rng.integers(2, 4))
    Acsr = sp.csr_matrix(A)
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(Acsr, np.int64)))
    ref = orc.dag_closure(Acsr, 64)
    assert info.events == 0 and ref.events == 0
    np.testing.assert_array_equal(sd.to_scipy(C).toarray(), ref.toarray(n))
    assert (info.cand_max, info.final_max) == (ref.cand_max, ref.final_max)


def test_dag_gpu_entry_point(orc):
    from redblackgraph_amd.gpu import CSRMatrixGPU, transitive_closure_dag_gpu
    A = synth.csr_pedigree(synth.PedigreeSpec(3000, 10, 0.8, 3), 4).astype(np.int32)
    Ag = CSRMatrixGPU.from_cpu(A)
    assert not Ag.triangular  # pedigrees are lower triangular
    C, diameter = transitive_closure_dag_gpu(Ag)
    ref = orc.dag_closure(A, 32)
    _assert_csr_equal(C.to_cpu(), sp.csr_matrix((ref.data, ref.indices, ref.indptr), shape=A.shape))
    assert diameter == ref.diameter_final
    e, d = transitive_closure_dag_gpu(CSRMatrixGPU.from_cpu(sp.csr_matrix((3, 3), dtype=np.int32)))
    assert e.nnz == 0 and d == 0
    D = sp.diags([1, -1, 1], 0, shape=(3, 3), dtype=np.int32, format="csr")
    np.testing.assert_array_equal(transitive_closure_dag_gpu(CSRMatrixGPU.from_cpu(D))[0].to_cpu().toarray(),
                                  D.toarray())
    from redblackgraph_amd.sparse.csgraph import CycleError
    with pytest.raises(CycleError):
        transitive_closure_dag_gpu(CSRMatrixGPU.from_cpu(sp.csr_matrix(np.array([[1, 2], [2, 1]], np.int32))))


# ---------------------------------------------------------------------------
# SpGEMM and squaring
# ---------------------------------------------------------------------------
def _random_rb(n, density, seed, dtype=np.int64):
    rng = np.random.default_rng(seed)
    A = sp.random(n, n, density=density, random_state=seed, format="csr",
                  data_rvs=lambda k: rng.integers(2, 12, k)).astype(dtype)
    A.setdiag(np.where(rng.random(n) < 0.5, -1, 1).astype(dtype))
    A = sp.csr_matrix(A)
    A.sort_indices()
    return A


@pytest.mark.parametrize("n,density,seed", [(50, 0.08, 1), (300, 0.02, 2), (200, 0.3, 3)])
def test_spgemm_vs_oracle(orc, n, density, seed):
    A = _random_rb(n, density, seed)
    B = _random_rb(n, density, seed + 10)
    Cp, Cj, Cx, ev = orc.spgemm(A, B)
    ref = sp.csr_matrix((Cx, Cj, Cp), shape=(n, n))
    if ev:
        pytest.skip("oracle overflow")
    C, events = sd.spgemm(sd.canonicalize(sd.from_scipy(A)), sd.canonicalize(sd.from_scipy(B)))
    _assert_csr_equal(sd.to_scipy(C), ref)


def test_spgemm_reference_analytic(golden):
    from redblackgraph_amd.gpu import CSRMatrixGPU, spgemm_upper_triangular
    from oracle.oracle import py_avos_product as P, py_avos_sum as S
    n = golden["spgemm_bidiagonal"]["n"]
    rows = np.concatenate([np.arange(n), np.arange(n - 1)])
    cols = np.concatenate([np.arange(n), np.arange(1, n)])
    data = np.concatenate([np.full(n, 2), np.full(n - 1, 3)]).astype(np.int32)
    A = sp.csr_matrix((data, (rows, cols)), shape=(n, n))
    C = spgemm_upper_triangular(CSRMatrixGPU.from_cpu(A, triangular=True)).to_cpu()
    assert np.array_equal(C.diagonal(0), np.full(n, P(2, 2)))
    assert np.array_equal(C.diagonal(1), np.full(n - 1, S(P(2, 3), P(3, 2))))
    assert np.array_equal(C.diagonal(2), np.full(n - 2, P(3, 3)))
    assert C.nnz == n + (n - 1) + (n - 2)
    n = golden["spgemm_diag"]["n"]
    d = np.arange(2, n + 2, dtype=np.int32)
    C = spgemm_upper_triangular(CSRMatrixGPU.from_cpu(sp.diags(d, 0, format="csr", dtype=np.int32),
                                                      triangular=True)).to_cpu()
    assert C.nnz == n
    assert np.array_equal(C.diagonal(), [P(int(x), int(x)) for x in d])
    I = sp.eye(n, dtype=np.int32, format="csr")
    assert np.array_equal(spgemm_upper_triangular(CSRMatrixGPU.from_cpu(I, triangular=True)).to_cpu().toarray(),
                          I.toarray())


def test_rb_matrix_matmul(orc, golden):
    from redblackgraph_amd.sparse import rb_matrix
    A = _random_rb(120, 0.05, 5)
    Cp, Cj, Cx, ev = orc.spgemm(A, A)
    _assert_csr_equal(rb_matrix(A) @ rb_matrix(A), sp.csr_matrix((Cx, Cj, Cp), shape=A.shape))
    g = golden["matmul_5"]
    for dt in [np.int32, np.uint32, np.int64, np.uint8]:
        M = np.array(g["A"]).astype(dt)
        R = rb_matrix(sp.csr_matrix(M))
        prod = (R @ R).toarray()
        np.testing.assert_array_equal(prod, np.array(g["AA"]).astype(dt))
        assert (R @ R).dtype == np.dtype(dt)


def _py_squaring(A):
    """The reference GPU test's CPU oracle (tests/gpu/test_transitive_closure.py:22-50):
    R <- R (+) R (x) R with plain AVOS sums until stable."""
    from oracle.oracle import py_avos_product as P, py_avos_sum as S
    R = np.array(A, dtype=np.int64)
    n = R.shape[0]
    for _ in range(64):
        R2 = np.array([[0] * n for _ in range(n)], dtype=np.int64)
        for i in range(n):
            for j in range(n):
                acc = 0
                for k in range(n):
                    acc = S(acc, P(int(R[i, k]), int(R[k, j])))
                R2[i, j] = acc
        Rn = np.array([[S(int(R[i, j]), int(R2[i, j])) for j in range(n)] for i in range(n)], dtype=np.int64)
        if np.array_equal(R, Rn):
            break
        R = Rn
    return R


@pytest.mark.parametrize("A", [
    [[1, 2, 0], [0, -1, 3], [0, 0, 1]],
    [[1, 2, 0, 0], [0, -1, 3, 0], [0, 0, 1, 2], [0, 0, 0, -1]],
    [[1, 2, 3, 0], [0, -1, 0, 2], [0, 0, 1, 3], [0, 0, 0, -1]],
    [[1, 2, 0, 0], [0, -1, 0, 0], [0, 0, 1, 3], [0, 0, 0, 1]],
])
def test_squaring_closure_small(A):
    from redblackgraph_amd.gpu import CSRMatrixGPU, transitive_closure_gpu
    A = np.array(A, dtype=np.int32)
    R, d = transitive_closure_gpu(CSRMatrixGPU.from_cpu(sp.csr_matrix(A), triangular=True))
    np.testing.assert_array_equal(R.to_cpu().toarray(), _py_squaring(A))
    R2, _ = CSRMatrixGPU.from_cpu(sp.csr_matrix(A), triangular=True).transitive_closure()
    np.testing.assert_array_equal(R2.to_cpu().toarray(), _py_squaring(A))


def test_squaring_matches_dag_on_pedigree(orc):
    from redblackgraph_amd.gpu import CSRMatrixGPU, transitive_closure_gpu, transitive_closure_dag_gpu
    A = synth.csr_pedigree(synth.PedigreeSpec(800, 8, 0.8, 3), 8).astype(np.int32)
    R, _ = transitive_closure_gpu(CSRMatrixGPU.from_cpu(A))
    D, _ = transitive_closure_dag_gpu(CSRMatrixGPU.from_cpu(A))
    _assert_csr_equal(R.to_cpu(), D.to_cpu())


def test_squaring_non_triangular(orc):
    """A DAG in a scrambled vertex order (neither triangle): squaring path."""
    from redblackgraph_amd.sparse.csgraph import transitive_closure_squaring
    A = synth.dense_pedigree(synth.PedigreeSpec(120, 6, 0.8, 3), 2)
    perm = np.random.default_rng(1).permutation(120)
    P = A[np.ix_(perm, perm)]
    res = transitive_closure_squaring(sp.csr_matrix(P))
    ref = orc.warshall(P)
    np.testing.assert_array_equal(res.W.toarray(), ref.W)


# ---------------------------------------------------------------------------
# CSRMatrixGPU plumbing
# ---------------------------------------------------------------------------
def test_csr_gpu_plumbing():
    from redblackgraph_amd.gpu import CSRMatrixGPU
    A = sp.csr_matrix(np.array([[1, 2, 0], [0, -1, 3], [0, 0, 1]], dtype=np.int32))
    g = CSRMatrixGPU.from_cpu(A)
    assert g.triangular and g.nnz == 5 and g.shape == (3, 3)
    _assert_csr_equal(g.to_cpu(), A)
    assert "CSRMatrixGPU(3x3, nnz=5 (upper triangular))" == repr(g)
    assert g.memory_usage()["total"] > 0
    with pytest.raises(ValueError):
        CSRMatrixGPU.from_cpu(sp.csr_matrix(np.array([[1, 0], [2, 1]], dtype=np.int32)), triangular=True)
    with pytest.raises(ValueError):
        CSRMatrixGPU(np.array([1, 2], np.int32), np.array([0, 5], np.int32), np.array([0, 1, 2]), (2, 2))
    with pytest.raises(ValueError):
        CSRMatrixGPU(np.array([1, 2], np.int32), np.array([0, 1], np.int32), np.array([0, 2, 1]), (2, 2))
    z = CSRMatrixGPU(np.array([1, 0, 3], np.int32), np.array([0, 1, 1], np.int32), np.array([0, 2, 3]), (2, 2))
    z.eliminate_zeros()
    assert z.nnz == 2
    c = g.copy()
    assert c.nnz == g.nnz and c.data.data_ptr() != g.data.data_ptr()
    prod = (g @ g).to_cpu().toarray()
    from oracle.oracle import py_mat_avos
    np.testing.assert_array_equal(prod, np.triu(py_mat_avos(A.toarray(), A.toarray())))


@pytest.mark.parametrize("seed", [0, 1])
def test_sparse_edge_insert_equals_reclosure(orc, seed):
    """The sparse edge relational composition on a resident closure equals
    closing the edited graph from scratch; an edge closing a cycle raises."""
    spec = synth.PedigreeSpec(3000, 12, 0.8, 3)
    A = synth.csr_pedigree(spec, seed)
    C, _ = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    rng = np.random.default_rng(seed)
    diag = A.diagonal()
    B = A.tolil()
    for _ in range(5):
        a = int(rng.integers(1500, 3000))
        rel = int(rng.choice([2, 3]))
        cand = np.nonzero(diag[:a] == (-1 if rel == 2 else 1))[0]
        b = int(rng.choice(cand))
        C = sd.edge_insert(C, a, b, rel)
        B[a, b] = rel
    ref = orc.dag_closure(B.tocsr(), 64)
    np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr)
    np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices)
    np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data)
    # an ancestor pointing back at its descendant
    rows = np.repeat(np.arange(3000), np.diff(ref.indptr))
    k = int(np.nonzero(ref.indices < rows)[0][0])
    with pytest.raises(ValueError, match="cycle"):
        sd.edge_insert(C, int(ref.indices[k]), int(rows[k]), 2)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 4])
def test_sparse_edge_insert_many_dense_and_sparse_forms(orc, seed, monkeypatch):
    """Several parent edges into a resident sparse closure: the dense-form
    batch (edge_insert_many) and the edge-by-edge sparse form both equal the
    oracle's closure of the edited graph; an edge closing a cycle raises."""
    spec = synth.PedigreeSpec(2500, 11, 0.8, 3)
    A = synth.csr_pedigree(spec, seed)
    C0, _ = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    rng = np.random.default_rng(100 + seed)
    diag = A.diagonal()
    B = A.tolil()
    edges = []
    for _ in range(12):
        a = int(rng.integers(1200, 2500))
        rel = int(rng.choice([2, 3]))
        cand = np.nonzero(diag[:a] == (-1 if rel == 2 else 1))[0]
        b = int(rng.choice(cand))
        edges.append((a, b, rel))
        B[a, b] = rel
    ref = orc.dag_closure(B.tocsr(), 64)
    got = {"dense": sd.edge_insert_many(C0, edges)}
    monkeypatch.setattr(sd, "DENSE_INSERT_MAX_BYTES", 0)
    got["sparse"] = sd.edge_insert_many(C0, edges)
    for form, C in got.items():
        np.testing.assert_array_equal(C.indptr.cpu().numpy(), ref.indptr, err_msg=form)
        np.testing.assert_array_equal(C.indices.cpu().numpy(), ref.indices, err_msg=form)
        np.testing.assert_array_equal(C.data.cpu().numpy(), ref.data, err_msg=form)
    monkeypatch.setattr(sd, "DENSE_INSERT_MAX_BYTES", 4 << 30)
    rows = np.repeat(np.arange(2500), np.diff(ref.indptr))
    k = int(np.nonzero(ref.indices < rows)[0][0])
    with pytest.raises(ValueError, match="cycle"):
        sd.edge_insert_many(got["dense"], [edges[0], (int(ref.indices[k]), int(rows[k]), 2)])
