"""Pin the CPU oracle: reference known-answer vectors, the reference's own
compiled C++ arithmetic (oracle/_ref, when built), closed forms and
independent genealogical checkers.  CPU only."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from tests.helpers import ahnentafel_closure, first_discovery_diameter, random_valid_pedigree, dense_triple_loop

DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def _as(dt, v):
    return np.array(v).astype(dt) if np.dtype(dt).kind == "u" else np.array(v, dtype=dt)


# ---------------------------------------------------------------------------
# scalar arithmetic
# ---------------------------------------------------------------------------
def test_product_table_all_widths(orc, golden):
    g = golden["product_table"]
    r = range(g["range"][0], g["range"][1] + 1)
    for dt in DTYPES:
        vals = iter(g["values"])
        for x, y in itertools.product(r, r):
            exp = next(vals)
            got, ev = orc.avos_product(int(_as(dt, x)), int(_as(dt, y)), dt)
            assert got == int(_as(dt, exp)), (dt, x, y)
            assert ev == 0


def test_sum_table_all_widths(orc, golden):
    g = golden["sum_table"]
    r = range(g["range"][0], g["range"][1] + 1)
    for dt in DTYPES:
        vals = iter(g["values"])
        for x, y in itertools.product(r, r):
            exp = next(vals)
            assert orc.avos_sum(int(_as(dt, x)), int(_as(dt, y)), dt) == int(_as(dt, exp))


def test_python_restatement_table(orc, golden):
    g = golden["product_table"]
    r = range(g["range"][0], g["range"][1] + 1)
    vals = iter(g["values"])
    for x, y in itertools.product(r, r):
        assert orc.py_avos_product(x, y) == next(vals)


def test_product_64bit_edge_and_overflow(orc, golden):
    for a, b, exp in golden["product_edge_64"]["cases"]:
        assert orc.avos_product_unsigned(a, b) == (exp, 0)
    for a, b in golden["product_overflow_64"]["cases"]:
        _, ev = orc.avos_product_unsigned(a, b)
        assert ev != 0


def test_closed_form(orc):
    # x (x) y = ((x - 1) << MSB(y)) + y for x >= 1, y >= 2 (Ahnentafel composition)
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x = int(rng.integers(1, 1 << 20))
        y = int(rng.integers(2, 1 << 20))
        m = y.bit_length() - 1
        got, ev = orc.avos_product(x, y, np.int64)
        assert got == ((x - 1) << m) + y and ev == 0


def test_oracle_matches_reference_cpp(orc):
    """Compare against the reference's own rbm_math.h compiled from its tree."""
    ref = orc.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(7)
    vals = list(range(-1, 40)) + [int(v) for v in rng.integers(0, 1 << 30, 300)]
    for a in vals:
        for b in vals[:60]:
            got, ev = orc.avos_product(a, b, np.int64)
            assert ev == 0
            assert got == np.array([ref.ref_avos_product_ll(a, b)], dtype=np.uint64).astype(np.int64)[0]
            assert orc.avos_sum(a, b, np.int64) == ref.ref_avos_sum_ll(a, b)
            if a < (1 << 20) and b < (1 << 10):
                g32, ev32 = orc.avos_product(a, b, np.int32)
                if ev32 == 0:
                    assert g32 == np.array([ref.ref_avos_product_i(a, b)], dtype=np.uint32).astype(np.int32)[0]
    with pytest.raises(OverflowError):
        ref.ref_avos_product_ll(4294967296, 4294967296)


# ---------------------------------------------------------------------------
# dense closure
# ---------------------------------------------------------------------------
def test_warshall_apsp5_all_dtypes(orc, golden):
    g = golden["apsp_5"]
    for dt in DTYPES:
        r = orc.warshall(_as(dt, g["A"]))
        np.testing.assert_array_equal(r.W, _as(dt, g["W"]))
        assert r.diameter == g["diameter"]
        assert r.report["n_cycle"] == 0 and r.report["n_overflow"] == 0


def test_warshall_apsp7(orc, golden):
    g = golden["apsp_7"]
    for dt in (np.int32, np.uint32):
        A = sp.coo_matrix((_as(dt, g["data"]), (g["rows"], g["cols"])), shape=(7, 7)).toarray()
        r = orc.warshall(A)
        np.testing.assert_array_equal(r.W.astype(np.int64), _as(dt, g["W"]).astype(np.int64))
        assert r.diameter == g["diameter"]
        c = orc.cy_floyd_warshall(A.astype(np.int32))
        np.testing.assert_array_equal(c.W, np.array(g["W"], dtype=np.int32))
        assert c.diameter == g["diameter"]


def test_warshall_cycle(orc, golden):
    A = np.array(golden["cycle_5"]["A"], dtype=np.int32)
    r = orc.warshall(A)
    assert r.report["n_cycle"] > 0
    c = orc.cy_floyd_warshall(A)
    assert c.report["raised"]
    # first detection of the naive loop is the Cython raise point
    assert (c.report["first_k"], c.report["first_i"]) == (r.report["first_k"], r.report["first_i"])


@pytest.mark.parametrize("seed", range(6))
def test_warshall_ahnentafel_and_diameter(orc, seed):
    rng = np.random.default_rng(100 + seed)
    A = random_valid_pedigree(11, rng)
    r = orc.warshall(A.astype(np.int32))
    np.testing.assert_array_equal(r.W, ahnentafel_closure(A))
    assert r.diameter == first_discovery_diameter(A)
    # idempotence and fixed point
    r2 = orc.warshall(r.W)
    np.testing.assert_array_equal(r2.W, r.W)
    WW, _ = orc.matmul(r.W, r.W)
    np.testing.assert_array_equal(np.where(r.W == 0, WW, np.minimum(np.where(WW == 0, r.W, WW), r.W)), r.W)


def test_warshall_ahnentafel_generated(orc):
    from redblackgraph_amd import synth
    A = synth.dense_pedigree(synth.PedigreeSpec(160, 8, 0.8, 3), seed=3)
    r = orc.warshall(A)
    np.testing.assert_array_equal(r.W, ahnentafel_closure(A))


def test_diameter_pin_catches_final_max(orc):
    """Pedigree collapse where the first-found value of a cell exceeds its final
    value: the trajectory rule (MSB 3) and a final-value rule (MSB 2) differ,
    so a diameter taken from the closed matrix fails here."""
    A = np.zeros((5, 5), dtype=np.int32)
    np.fill_diagonal(A, [-1, -1, -1, 1, -1])
    A[1, 0] = 2  # 1's father is 0
    A[2, 1] = 2  # 2's father is 1
    A[3, 0] = 2  # 3 (female)'s father is 0
    A[4, 2] = 2  # 4's father is 2: 0 is found first as 8 (via 2, 1) ...
    A[4, 3] = 3  # ... then as 6 via mother 3
    r = orc.warshall(A)
    assert r.W[4, 0] == 6
    assert first_discovery_diameter(A) == 3
    assert r.diameter == 3
    assert int(r.W.max()).bit_length() - 1 == 2


# ---------------------------------------------------------------------------
# DAG closure, topological sort, SpGEMM, matmul, composition
# ---------------------------------------------------------------------------
def test_dag_known_answers(orc, golden):
    g = golden
    for width in (32, 64):
        r = orc.dag_closure(sp.csr_matrix(np.array(g["dag_chain"]["A"])), width)
        assert r.toarray(3)[0, 2] == g["dag_chain"]["W02"]
        r = orc.dag_closure(sp.csr_matrix(np.array(g["dag_known_4"]["A"])), width)
        np.testing.assert_array_equal(r.toarray(4), g["dag_known_4"]["W"])
        r = orc.dag_closure(sp.csr_matrix(np.array(g["dag_diamond"]["A"])), width)
        assert r.toarray(4)[0, 3] == g["dag_diamond"]["W03"]
        r = orc.dag_closure(sp.csr_matrix(np.array(g["dag_parity"]["A"])), width)
        np.testing.assert_array_equal(np.diag(r.toarray(3)), g["dag_parity"]["diag"])
        r = orc.dag_closure(sp.csr_matrix(np.array(g["dag_cycle"]["A"])), width)
        assert r.cycle_vertex >= 0


def test_dag_equals_fw_on_reference_seeds(orc, golden):
    from redblackgraph_amd.synth import random_upper_dag
    for seed, n in golden["dag_vs_fw_seeds"]["cases"]:
        A = random_upper_dag(n, seed)
        fw = orc.cy_floyd_warshall(A)
        dag = orc.dag_closure(sp.csr_matrix(A), 32)
        np.testing.assert_array_equal(dag.toarray(n), fw.W)


@pytest.mark.parametrize("seed", range(4))
def test_dag_ahnentafel(orc, seed):
    rng = np.random.default_rng(300 + seed)
    A = random_valid_pedigree(40, rng)
    r = orc.dag_closure(sp.csr_matrix(A), 64)
    np.testing.assert_array_equal(r.toarray(40), ahnentafel_closure(A))


def _candidate_max_by_ahnentafel(A: np.ndarray) -> int:
    """max |candidate| of the DAG closure, from the pedigree's meaning: the
    direct entries, and for each child -> parent edge (v, w) of value a and
    every ancestor x of w with Ahnentafel number y = C[w][x] (C from the BFS
    closure), the number of x seen from v through w: ((a - 1) << MSB(y)) + y
    for y >= 2 (a prefix of the path digits), and a itself for x = w (the
    parity filter keeps a: fathers are red, mothers black)."""
    C = ahnentafel_closure(A)
    best = int(np.abs(A).max()) if A.size else 0
    n = A.shape[0]
    for v in range(n):
        for w in np.nonzero(A[v])[0]:
            if w == v:
                continue
            a = int(A[v, w])
            for x in np.nonzero(C[w])[0]:
                y = int(C[w, x])
                if x == w:
                    cand = a
                else:
                    cand = ((a - 1) << (y.bit_length() - 1)) + y
                best = max(best, cand)
    return best


@pytest.mark.parametrize("seed", range(6))
def test_dag_candidate_max_pinned(orc, seed):
    """The Cython DAG closure's diameter rule (_transitive_closure_dag.pyx:
    261-265, 291-295, 347-349): MSB of the largest |candidate| over direct
    edges and every nonzero A[v][w] (x) C[w][x].  Checked against the
    candidates enumerated from the Ahnentafel closure on pedigrees with
    collapse (several paths to one ancestor), where the largest candidate is
    a longer path's number and exceeds the largest final value."""
    rng = np.random.default_rng(900 + seed)
    A = random_valid_pedigree(48, rng, p=0.9, lookback=3)
    r = orc.dag_closure(sp.csr_matrix(A), 64)
    exp = _candidate_max_by_ahnentafel(A)
    assert r.cand_max == exp
    assert r.final_max == int(ahnentafel_closure(A).max())


def test_dag_candidate_max_differs_from_final():
    """At least one seeded case separates the two rules in the diameter
    itself, so a closure that reports MSB(max final) fails the pin above."""
    diffs = 0
    for seed in range(6):
        rng = np.random.default_rng(900 + seed)
        A = random_valid_pedigree(48, rng, p=0.9, lookback=3)
        cand = _candidate_max_by_ahnentafel(A)
        fin = int(ahnentafel_closure(A).max())
        assert cand >= fin
        if cand.bit_length() != fin.bit_length():
            diffs += 1
    assert diffs > 0


def test_topological_sort(orc):
    from redblackgraph_amd.synth import random_upper_dag
    A = sp.csr_matrix(random_upper_dag(30, 5))
    order, bad = orc.topological_sort(A)
    assert bad == -1
    pos = np.empty(30, int)
    pos[order] = np.arange(30)
    r, c = A.nonzero()
    assert np.all((r == c) | (pos[r] < pos[c]))
    _, bad = orc.topological_sort(sp.csr_matrix(np.array([[1, 2], [2, 1]])))
    assert bad >= 0


def test_spgemm_matches_reference_cpp_and_triple_loop(orc):
    rng = np.random.default_rng(11)
    A = sp.random(40, 40, density=0.12, random_state=3, format="csr", data_rvs=lambda k: rng.integers(2, 9, k))
    A.setdiag(np.where(rng.random(40) < 0.5, -1, 1))
    A = sp.csr_matrix(A.astype(np.int64))
    A.sort_indices()
    Cp, Cj, Cx, ev = orc.spgemm(A, A)
    C = sp.csr_matrix((Cx, Cj, Cp), shape=(40, 40)).toarray()
    D = dense_triple_loop(A.toarray(), A.toarray(), orc.py_avos_product, orc.py_avos_sum)
    np.testing.assert_array_equal(C, D)
    ref = orc.ref_lib()
    if ref is not None:
        Ap = A.indptr.astype(np.int32); Aj = A.indices.astype(np.int32); Ax = A.data.astype(np.int64)
        rp = np.zeros(41, np.int32)
        ref.ref_matmat_pass1(40, 40, Ap.ctypes.data, Aj.ctypes.data, Ap.ctypes.data, Aj.ctypes.data, rp.ctypes.data)
        rj = np.zeros(rp[-1], np.int32); rx = np.zeros(rp[-1], np.int64); rp2 = np.zeros(41, np.int32)
        ref.ref_matmat_pass2(40, 40, Ap.ctypes.data, Aj.ctypes.data, Ax.ctypes.data, Ap.ctypes.data, Aj.ctypes.data,
                             Ax.ctypes.data, rp2.ctypes.data, rj.ctypes.data, rx.ctypes.data)
        np.testing.assert_array_equal(rp2, Cp)
        np.testing.assert_array_equal(rj, Cj)  # same (unsorted) column order
        np.testing.assert_array_equal(rx, Cx)


def test_matmul_known(orc, golden):
    g = golden["matmul_5"]
    for dt in DTYPES:
        A = _as(dt, g["A"])
        AA, ev = orc.matmul(A, A)
        np.testing.assert_array_equal(AA, _as(dt, g["AA"]))
        S3, _ = orc.matmul(AA, A)
        np.testing.assert_array_equal(S3, _as(dt, g["A_star"]))
    A = np.array(g["A"], dtype=np.int64)
    np.testing.assert_array_equal(orc.matmul(A, A)[0], orc.py_mat_avos(A, A))


def test_compositions(orc, golden):
    for key in ("vertex_compose_1", "vertex_compose_2", "vertex_compose_14"):
        g = golden[key]
        out, rc, ev = orc.vertex_compose(g["u"], np.array(g["A"], np.int32), g["v"], g["color"])
        assert rc == 0
        np.testing.assert_array_equal(out, g["out"])
    g = golden["vertex_cycle"]
    W = orc.warshall(np.array(g["A"], np.int32)).W
    out, rc, ev = orc.vertex_compose(g["u"], W, g["v"], g["color"])
    idx = -1 - rc
    msg = f"Relational composition would result in a cycle. Idx: {idx}, u_i: {out[-1, idx]}, v_i: {out[idx, -1]}"
    assert msg == g["message"]
    for key in ("edge_compose_simple", "edge_compose_15"):
        g = golden[key]
        out, ev = orc.edge_compose(np.array(g["R"], np.int32), g["alpha"], g["beta"], g["rel"])
        np.testing.assert_array_equal(out, g["out"])


def test_canonical_ordering_golden(orc, golden):
    g = golden["canonical_ordering_7"]
    B, p, q = orc.canonical_ordering(g["A_star"])
    np.testing.assert_array_equal(p, g["permutation"])
    np.testing.assert_array_equal(B, g["canonical"])
    g2 = golden["ordering_two_components"]
    _, p2, q2 = orc.canonical_ordering(g2["A"])
    assert len(q2) == g2["n_components"] and sum(q2.values()) == g2["total"]
    assert sorted(p2.tolist()) == list(range(4))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_canonical_ordering_keys_brute_force(orc, seed):
    """Every adjacent pair of the oracle's order obeys the reference's key,
    recomputed entry by entry from the definition, and components sit in
    contiguous runs."""
    A = random_valid_pedigree(40, np.random.default_rng(seed))
    C = np.asarray(orc.warshall(A).W)
    B, p, q = orc.canonical_ordering(C)
    n = C.shape[0]
    comp = {}
    for c0 in range(n):  # union-find over row/column nonzeros
        comp.setdefault(c0, c0)
    def find(x):
        while comp[x] != x:
            x = comp[x]
        return x
    for i in range(n):
        for j in range(n):
            if C[i][j] != 0:
                a, b = find(i), find(j)
                comp[max(a, b)] = min(a, b)
    root = [find(i) for i in range(n)]
    ids = {r: k for k, r in enumerate(sorted(set(root)))}
    size = {r: root.count(r) for r in set(root)}
    def key(x):
        anc = sum(orc.py_msb(int(v)) for v in C[x] if v != 0)
        mr = max([0] + [int(C[j][x]) for j in range(n) if C[j][x] != 0])
        return (-size[root[x]], ids[root[x]], -anc, mr, -int(C[x][x]), x)
    keys = [key(int(x)) for x in p]
    assert keys == sorted(keys)
    np.testing.assert_array_equal(B, C[np.ix_(p, p)])
    assert sum(q.values()) == n


@pytest.mark.parametrize("seed", [3, 4])
def test_canonical_ordering_csr_equals_dense(orc, seed):
    """The vectorised CSR restatement (used as the timed CPU baseline) gives
    the dense restatement's permutation and component sizes, including values
    where floor(log2) sits next to a power of two."""
    rng = np.random.default_rng(seed)
    blocks = [np.asarray(orc.warshall(random_valid_pedigree(int(rng.integers(10, 40)), rng)).W) for _ in range(3)]
    C = sp.block_diag(blocks).toarray()
    C[0, C.shape[0] - 1] = (1 << 40) - 1  # joins two components; MSB 39, a float-rounding edge
    idx = rng.permutation(C.shape[0])
    C = C[np.ix_(idx, idx)]
    _, p_dense, q_dense = orc.canonical_ordering(C)
    p_csr, q_csr = orc.canonical_ordering_csr(sp.csr_matrix(C))
    np.testing.assert_array_equal(p_csr, p_dense)
    assert q_csr == q_dense
