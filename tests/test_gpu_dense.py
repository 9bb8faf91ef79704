"""GPU parity of the dense path (element-wise AVOS, Floyd-Warshall closure,
AVOS matrix product, relational composition) against the CPU oracle, the
reference's known answers and size-independent properties.  Every call goes
through librbg_hip.so's C ABI."""
import itertools

import numpy as np
import pytest

import redblackgraph_amd as rb
from helpers import overflow_chain
from redblackgraph_amd import dense, synth, _lib
from redblackgraph_amd._dtypes import elem_type

pytestmark = pytest.mark.gpu

DTYPES = [np.int8, np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64]


def _as(dt, v):
    return np.array(v).astype(dt) if np.dtype(dt).kind == "u" else np.array(v, dtype=dt)


def _fw(A, force_general=False):
    et = elem_type(A.dtype)
    from redblackgraph_amd._dtypes import widen, narrow
    W, out = dense.fw_closure(_lib.to_device(widen(A, et)), et, force_general=force_general)
    return narrow(W.cpu().numpy(), et), out


# ---------------------------------------------------------------------------
# element-wise kernels
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt", [np.int32, np.int64])
def test_elementwise_vs_oracle(orc, dt):
    from redblackgraph_amd.gpu import avos_sum_gpu, avos_product_gpu
    vals = list(range(-1, 70))
    x = np.array([a for a, b in itertools.product(vals, vals)], dtype=dt)
    y = np.array([b for a, b in itertools.product(vals, vals)], dtype=dt)
    zs = avos_sum_gpu(x, y)
    zp = avos_product_gpu(x, y)
    for i in range(0, len(x), 7):
        assert zs[i] == orc.avos_sum(int(x[i]), int(y[i]), dt)
        assert zp[i] == orc.avos_product(int(x[i]), int(y[i]), dt)[0]


def test_elementwise_table(golden):
    from redblackgraph_amd.gpu import avos_product_gpu, avos_sum_gpu
    g = golden["product_table"]
    r = list(range(g["range"][0], g["range"][1] + 1))
    x = np.array([a for a, b in itertools.product(r, r)], dtype=np.int32)
    y = np.array([b for a, b in itertools.product(r, r)], dtype=np.int32)
    np.testing.assert_array_equal(avos_product_gpu(x, y), g["values"])
    g = golden["sum_table"]
    r = list(range(g["range"][0], g["range"][1] + 1))
    x = np.array([a for a, b in itertools.product(r, r)], dtype=np.int32)
    y = np.array([b for a, b in itertools.product(r, r)], dtype=np.int32)
    np.testing.assert_array_equal(avos_sum_gpu(x, y), g["values"])


# ---------------------------------------------------------------------------
# Floyd-Warshall
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt", DTYPES)
def test_fw_apsp5(golden, dt):
    g = golden["apsp_5"]
    A = rb.array(g["A"], dtype=dt)
    res = A.transitive_closure()
    np.testing.assert_array_equal(res.W, _as(dt, g["W"]))
    assert res.diameter == g["diameter"]
    assert isinstance(res.W, rb.array)


def test_fw_apsp5_module_function(golden):
    res = rb.transitive_closure(rb.array(golden["apsp_5"]["A"]))
    np.testing.assert_array_equal(res.W, golden["apsp_5"]["W"])
    assert res.diameter == 3


def test_fw_apsp7(golden):
    import scipy.sparse as sp
    g = golden["apsp_7"]
    for dt in (np.int32, np.uint32):
        A = sp.coo_matrix((_as(dt, g["data"]), (g["rows"], g["cols"])), shape=(7, 7)).toarray()
        res = rb.array(A).transitive_closure()
        np.testing.assert_array_equal(res.W.astype(np.int64), _as(dt, g["W"]).astype(np.int64))
        assert res.diameter == g["diameter"]


def test_fw_cycle_message(orc, golden):
    A = np.array(golden["cycle_5"]["A"], dtype=np.int32)
    ref = orc.warshall(A).report
    with pytest.raises(ValueError) as e:
        rb.array(A).transitive_closure()
    msg = str(e.value)
    assert golden["cycle_5"]["substring"] in msg
    k, i = ref["last_k"], ref["last_i"]
    exp = (f"Error: cycle detected! Vertex {i} has a path to itself. "
           f"A({i},{k})={ref['last_a'] & 0xFFFFFFFF}, A({k},{i})={ref['last_b'] & 0xFFFFFFFF}")
    assert msg == exp


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_fw_dense512_config(orc, seed):
    """The N=512 int32 pedigree workload: W and diameter bit-exact vs the oracle."""
    A = synth.dense_pedigree(synth.DENSE_512, seed)
    ref = orc.warshall(A)
    W, out = _fw(A)
    assert out.mode == "fast"
    np.testing.assert_array_equal(W, ref.W)
    assert out.diameter == ref.diameter
    assert out.status.cycle_count == 0 and out.status.events == 0


@pytest.mark.parametrize("seed", [3, 4])
def test_fw_general_equals_fast(orc, seed):
    A = synth.dense_pedigree(synth.PedigreeSpec(384, 12, 0.8, 4), seed)
    Wf, of = _fw(A)
    Wg, og = _fw(A, force_general=True)
    assert of.mode == "fast" and og.mode == "general"
    np.testing.assert_array_equal(Wf, Wg)
    assert of.diameter == og.diameter
    ref = orc.warshall(A)
    np.testing.assert_array_equal(Wf, ref.W)
    assert of.diameter == ref.diameter


@pytest.mark.parametrize("seed,n", [(42, 20), (123, 15), (456, 15), (7, 200)])
def test_fw_reference_random_dags(orc, seed, n):
    """The reference's parity-inconsistent random DAGs (off-diagonal 1s): general mode."""
    A = synth.random_upper_dag(n, seed)
    ref = orc.warshall(A)
    W, out = _fw(A)
    assert out.mode == "general"
    np.testing.assert_array_equal(W, ref.W)
    assert out.diameter == ref.diameter


def test_fw_int64(orc):
    A = synth.dense_pedigree(synth.PedigreeSpec(300, 14, 0.75, 4), 9, dtype=np.int64)
    ref = orc.warshall(A)
    W, out = _fw(A)
    assert out.mode == "fast"
    np.testing.assert_array_equal(W, ref.W)
    assert out.diameter == ref.diameter


def test_fw_cyclic_pedigree_errors(orc):
    """A pedigree with back edges: last (core) and first (csgraph) detections
    and their operands must equal the naive loop's."""
    A = synth.dense_pedigree(synth.PedigreeSpec(260, 10, 0.8, 3), 5)
    rng = np.random.default_rng(5)
    for _ in range(6):
        i, j = sorted(rng.integers(0, 260, 2))
        if i != j:
            A[i, j] = 2 if A[j, j] == -1 else 3
    ref = orc.warshall(A).report
    assert ref["n_cycle"] > 0
    W, out = _fw(A)
    s = out.status
    assert s.cycle_count > 0
    assert out.key_kij(s.cycle_first_key)[:2] == (ref["first_k"], ref["first_i"])
    assert out.key_kij(s.cycle_last_key)[:2] == (ref["last_k"], ref["last_i"])
    assert (s.cycle_first_a, s.cycle_first_b) == (ref["first_a"], ref["first_b"])
    assert (s.cycle_last_a, s.cycle_last_b) == (ref["last_a"], ref["last_b"])
    from redblackgraph_amd.sparse.csgraph import CycleError
    with pytest.raises(ValueError):
        rb.array(A).transitive_closure()
    with pytest.raises(CycleError) as e:
        rb.sparse.transitive_closure_floyd_warshall(rb.array(A))
    assert e.value.vertex == ref["first_i"]


def _pivot_cases():
    # vertex 1 has an absent diagonal and a 2-cycle with vertex 0: the naive
    # loop's last detection is the pivot cell (k=1, i=1), W[1][1] (x) W[1][1]
    two = np.zeros((6, 6), dtype=np.int32)
    np.fill_diagonal(two, [-1, 0, 1, -1, 1, -1])
    two[0, 1], two[1, 0], two[2, 3] = 2, 3, 2
    # a diagonal value >= 2 in the input is a cycle at the very first step
    diag = synth.dense_pedigree(synth.PedigreeSpec(150, 6, 0.8, 2), 3)
    diag[7, 7] = 2
    return {"two_cycle_absent_diag": two, "diag_two": diag}


@pytest.mark.parametrize("case", ["two_cycle_absent_diag", "diag_two"])
def test_fw_pivot_cell_detections(orc, case):
    """Detections made by products through the pivot cell (k, k): the fast
    kernels skip that row and column, so these inputs must come out of the
    general kernels with every event of the naive loop."""
    A = _pivot_cases()[case]
    ref = orc.warshall(A)
    rep = ref.report
    assert rep["n_cycle"] > 0
    W, out = _fw(A)
    s = out.status
    assert out.mode == "general"
    np.testing.assert_array_equal(W, ref.W)
    assert s.cycle_count == rep["n_cycle"]
    assert out.key_kij(s.cycle_first_key)[:2] == (rep["first_k"], rep["first_i"])
    assert out.key_kij(s.cycle_last_key)[:2] == (rep["last_k"], rep["last_i"])
    assert (s.cycle_first_a, s.cycle_first_b) == (rep["first_a"], rep["first_b"])
    assert (s.cycle_last_a, s.cycle_last_b) == (rep["last_a"], rep["last_b"])
    if case == "two_cycle_absent_diag":
        assert (rep["last_k"], rep["last_i"]) == (1, 1)


def test_fw_1024_vs_oracle(orc):
    A = synth.dense_pedigree(synth.PedigreeSpec(1024, 16, 0.75, 4), 11)
    ref = orc.warshall(A)
    W, out = _fw(A)
    np.testing.assert_array_equal(W, ref.W)
    assert out.diameter == ref.diameter


@pytest.mark.parametrize("dt,n,reverse", [(np.int32, 200, False), (np.int32, 512, True), (np.int32, 1100, False),
                                          (np.int32, 1100, True), (np.int64, 300, False), (np.int64, 700, True)])
def test_fw_triangular_schedule_vs_oracle(orc, dt, n, reverse):
    """Lower (parents first) and upper (numbered backwards) triangular
    pedigrees take the triangular schedule; W and the trajectory diameter
    equal the oracle's naive loop and the full schedule's."""
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 17 + n, dtype=dt)
    if reverse:
        A = synth.reverse_vertices(A)
    ref = orc.warshall(A)
    W, out = _fw(A)
    assert out.mode == "fast" and out.schedule == ("upper" if reverse else "lower")
    np.testing.assert_array_equal(W, ref.W)
    assert out.diameter == ref.diameter
    W2, out2 = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype), schedule="full")
    assert out2.schedule == "full"
    np.testing.assert_array_equal(W2.cpu().numpy(), ref.W)
    assert out2.diameter == ref.diameter


@pytest.mark.parametrize("dt,n", [(np.int32, 128), (np.int32, 700), (np.int32, 2100), (np.int32, 3000),
                                  (np.int64, 128), (np.int64, 300), (np.int64, 900), (np.int64, 2000),
                                  (np.int64, 3000)])
def test_fw_persistent_closure_vs_steps_and_oracle(orc, dt, n):
    """The persistent lower triangular closure (one launch, work items waiting
    on each other) against the same schedule launched step by step and the
    oracle: W and the trajectory diameter, three runs each (the items' order
    on the GPU differs from run to run), no wait running out."""
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 29 + n, dtype=dt)
    ref = orc.warshall(A) if n <= 1100 else None
    Ws, outs = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype), schedule="steps")
    assert outs.schedule == "lower-steps"
    Ws = Ws.cpu().numpy()
    if ref is not None:
        np.testing.assert_array_equal(Ws, ref.W)
        assert outs.diameter == ref.diameter
    for _ in range(3):
        W, out = _fw(A)
        assert out.mode == "fast" and out.schedule == "lower"
        assert not out.status.flags & (_lib.F_STALLED | _lib.F_NEED_GENERAL)
        np.testing.assert_array_equal(W, Ws)
        assert out.diameter == outs.diameter


@pytest.mark.parametrize("dt,n,lda,bad", [(np.int32, 300, 300, False), (np.int32, 261, 300, False),
                                          (np.int64, 130, 133, False), (np.int32, 200, 200, True)])
def test_fw_load_equals_copy_precheck_and_encoding(dt, n, lda, bad):
    """rbg_fw_load_* (the fast path's one-pass copy-in) against what it
    replaces: a zero-padded copy, rbg_fw_precheck_*'s flags, the encoding
    (absent -> all ones, -1 -> 0, v >= 1 -> v) and the largest input value in
    aux[7]; then a closure through it against one through a plain copy.  Row
    strides wider than n and orders that are not multiples of four take the
    scalar loads."""
    t = _lib.torch()
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 7 + n, dtype=dt)
    if bad:
        A = A.copy()
        A[5, 3] = 1  # an off-diagonal 1: outside the fast domain
    big = np.zeros((n, lda), dtype=dt)
    big[:, :n] = A
    Bd = _lib.to_device(big)[:, :n]  # a view with row stride lda
    npad = dense.padded_order(n)
    sfx = "i32" if dt == np.int32 else "i64"
    st = dense.status()
    st.reset()
    Wl = t.empty((npad, npad), dtype=Bd.dtype, device=Bd.device)
    _lib.call(f"rbg_fw_load_{sfx}", _lib.ptr(Bd), n, lda, _lib.ptr(Wl), npad, st.p, _lib.stream_handle(None))
    s1 = st.read()
    P = t.zeros((npad, npad), dtype=Bd.dtype, device=Bd.device)
    P[:n, :n] = _lib.to_device(A)
    st.reset()
    _lib.call(f"rbg_fw_precheck_{sfx}", _lib.ptr(P), npad, npad, st.p, _lib.stream_handle(None))
    s2 = st.read()
    assert s1.flags == s2.flags
    assert bool(s1.flags & _lib.F_PRECOND_FAIL) == bad
    assert s1.aux[7] == int(A.max())
    if bad:
        return
    # the encoding: absent (0) -> all ones, -1 -> 0, v >= 1 -> v
    enc = P.clone()
    enc[P == 0] = -1
    enc[P == -1] = 0
    assert bool(t.equal(Wl, enc))
    # a closure through the fused load equals one through a plain copy
    W1, o1 = dense.fw_closure(Bd, elem_type(A.dtype))
    W2, o2 = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype))
    assert bool(t.equal(W1, W2)) and o1.status.diameter_max == o2.status.diameter_max
    assert o1.status.flags == 0 and o1.status.aux[7] == 0


def test_fw_closure_of_views():
    """Input views the one-pass load cannot read with a row stride (a
    transposed view, a broadcast row) close like their contiguous copies."""
    t = _lib.torch()
    et = elem_type(np.int32)
    A = synth.dense_pedigree(synth.component_spec(300, 0.8, 4), 11)
    Ad = _lib.to_device(A)
    Wt, ot = dense.fw_closure(Ad.t(), et)
    W2, o2 = dense.fw_closure(Ad.t().contiguous(), et)
    assert bool(t.equal(Wt, W2)) and ot.diameter == o2.diameter and ot.schedule == "upper"
    row = t.zeros(5, dtype=t.int32, device=Ad.device)
    B = row.unsqueeze(0).expand(5, 5)  # row stride 0: all absent
    Wb, ob = dense.fw_closure(B, et)
    assert bool(t.equal(Wb, t.zeros((5, 5), dtype=t.int32, device=Ad.device)))


@pytest.mark.parametrize("dt,n", [(np.int32, 2100), (np.int64, 900)])
def test_fw_persistent_wait_limit_reruns_steps(dt, n):
    """A wait of the persistent closure that runs out, forced from the host
    with a one-tick limit: every workgroup ends, the closure is flagged
    stalled, and fw_closure reruns the per-step schedule from the input --
    the same W, diameter and status as an unforced closure."""
    L = _lib.load()
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 29 + n, dtype=dt)
    W0, out0 = _fw(A)
    assert out0.schedule == "lower"
    prev = L.rbg_fw_flow_wait_limit(1)
    try:
        W1, out1 = _fw(A)
    finally:
        L.rbg_fw_flow_wait_limit(prev)
    assert out1.mode == "fast" and out1.schedule == "lower-steps"
    np.testing.assert_array_equal(W1, W0)
    assert out1.diameter == out0.diameter and out1.status.flags == out0.status.flags == 0


@pytest.mark.parametrize("dt,n,wrong", [(np.int32, 640, 2), (np.int32, 640, 1), (np.int64, 384, 2)])
def test_fw_triangular_wrong_side_detected(dt, n, wrong):
    """rbg_fw_dense_tri_* told the wrong side (tri = 2 for a lower triangular
    pedigree, or tri = 1 once one entry above the diagonal is added) stops
    the fast run with RBG_F_NEED_GENERAL instead of returning a partial
    closure; the right side on the same input returns the oracle's W."""
    t = _lib.torch()
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 5 + n, dtype=dt)
    if wrong == 1:
        A = A.copy()
        A[3, n - 7] = 2
    sfx = "i32" if dt == np.int32 else "i64"
    esz = np.dtype(dt).itemsize
    dW = _lib.to_device(A)
    ws = t.empty(int(_lib.load().rbg_fw_tri_workspace_bytes(n, esz)), dtype=t.uint8, device=dW.device)
    st = dense.status()
    st.reset()
    _lib.call(f"rbg_fw_dense_tri_{sfx}", _lib.ptr(dW), n, n, 8 * esz, 0, 0, wrong, _lib.ptr(ws), ws.numel(), st.p,
              _lib.stream_handle(None))
    assert st.read().flags & _lib.F_NEED_GENERAL
    if wrong == 2:
        from oracle import oracle as orc
        dW = _lib.to_device(A)
        st.reset()
        _lib.call(f"rbg_fw_dense_tri_{sfx}", _lib.ptr(dW), n, n, 8 * esz, 0, 0, 1, _lib.ptr(ws), ws.numel(), st.p,
                  _lib.stream_handle(None))
        s = st.read()
        assert not s.flags & _lib.F_NEED_GENERAL
        ref = orc.warshall(A)
        np.testing.assert_array_equal(dW.cpu().numpy(), ref.W)
        assert int(s.diameter_max).bit_length() - 1 == ref.diameter


def _digests():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "fw_digests.json")
    with open(path) as f:
        return json.load(f)


def _workload(name):
    """The matrix a digest in tests/golden/fw_digests.json was made from."""
    import importlib.util
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "make_fw_digests.py")
    spec = importlib.util.spec_from_file_location("make_fw_digests", path)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.workload_matrix(name), mod.digest


def _check_digest(name, W, out):
    import hashlib
    ref = _digests()[name]
    assert np.dtype(W.dtype).name == ref["dtype"]
    got = hashlib.sha256(np.ascontiguousarray(W).tobytes()).hexdigest()
    assert int(np.count_nonzero(W)) == ref["nnz_W"]
    assert int(W.max()) == ref["max_W"]
    assert got == ref["sha256_W"]
    assert out.diameter == ref["diameter"]
    assert out.status.cycle_count == 0 and out.status.events == 0
    assert not out.status.flags & _lib.F_STALLED


@pytest.mark.parametrize("name,schedule", [
    ("dense2304_i32", "lower"),            # parents first: lower triangular
    ("dense2304_reversed_i32", "upper"),   # numbered backwards: upper triangular
    ("dense2304_shuffled_i32", "full"),    # renumbered at random: neither
])
def test_fw_2304_vs_oracle(name, schedule):
    """N=2304 int32 (above 2048, so phase 3 runs the 128 x 128 tile kernel
    k_fw_tile_fast<int> of the N=8192 workload) against the oracle's naive
    k-i-j loop, run in full when the digest was made: W bit-exact (sha256)
    and the trajectory diameter (warshall.c.in:89,106).  Each input also runs
    the full schedule, and the first one general mode."""
    A, _ = _workload(name)
    W, out = _fw(A)
    assert out.mode == "fast" and out.schedule == schedule
    _check_digest(name, W, out)
    W2, out2 = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype), schedule="full")
    assert out2.schedule == "full"
    _check_digest(name, W2[:A.shape[0], :A.shape[0]].cpu().numpy(), out2)
    if schedule == "lower":  # the persistent closure above; one launch per pivot step here
        W3, out3 = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype), schedule="steps")
        assert out3.schedule == "lower-steps"
        _check_digest(name, W3[:A.shape[0], :A.shape[0]].cpu().numpy(), out3)
    if name == "dense2304_i32":
        Wg, og = _fw(A, force_general=True)
        assert og.mode == "general"
        _check_digest(name, Wg, og)


@pytest.mark.parametrize("name", ["dense8192_i32", "dense4096_i64"])
def test_fw_full_size_vs_oracle(name):
    """The full-size dense workloads (N=8192 int32, N=4096 int64) against the
    oracle's naive loop run once in full on the CPU: W digest and trajectory
    diameter, for the triangular schedule the closure takes, the full
    schedule, and general mode (the full arithmetic)."""
    if name not in _digests():
        pytest.skip(f"oracle digest {name} not generated")
    A, _ = _workload(name)
    W, out = _fw(A)
    assert out.mode == "fast" and out.schedule == "lower"
    _check_digest(name, W, out)
    del W
    W3, out3 = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype), schedule="steps")
    assert out3.schedule == "lower-steps"
    _check_digest(name, W3.cpu().numpy(), out3)
    del W3
    W2, out2 = dense.fw_closure(_lib.to_device(A), elem_type(A.dtype), schedule="full")
    _check_digest(name, W2.cpu().numpy(), out2)
    del W2
    Wg, og = _fw(A, force_general=True)
    assert og.mode == "general"
    _check_digest(name, Wg, og)


def test_fw_8192_properties():
    """Full-size workload: fast vs general arithmetic agree, the closure is a
    fixed point (closing it again changes nothing), lower-triangular input
    stays lower triangular."""
    t = _lib.torch()
    A = synth.dense_pedigree(synth.DENSE_8192, 0)
    et = elem_type(A.dtype)
    Ad = _lib.to_device(A)
    W1, o1 = dense.fw_closure(Ad, et)
    assert o1.mode == "fast"
    W2, o2 = dense.fw_closure(W1, et)
    assert bool(t.equal(W1, W2))
    assert bool((t.triu(W1, diagonal=1) == 0).all())
    diag = t.diagonal(W1)
    assert bool(t.equal(diag, t.diagonal(Ad)))
    # spot-check rows against the Ahnentafel rule: W[i][p] for each direct parent p
    Wc = W1.cpu().numpy()
    rows, cols = np.nonzero(np.tril(A, -1))
    np.testing.assert_array_equal(Wc[rows, cols], A[rows, cols])
    assert o1.diameter >= 1


# ---------------------------------------------------------------------------
# AVOS product and composition
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("dt", DTYPES)
def test_matmul_known(golden, dt):
    g = golden["matmul_5"]
    A = rb.array(g["A"], dtype=dt)
    np.testing.assert_array_equal(A @ A, _as(dt, g["AA"]))
    np.testing.assert_array_equal(A @ A @ A, _as(dt, g["A_star"]))
    np.testing.assert_array_equal((A @ A) @ A, _as(dt, g["A_star"]))
    u = rb.array(g["u"], dtype=dt).reshape((1, 5))
    v = rb.array(g["v"], dtype=dt).reshape((5, 1))
    S = rb.array(g["A_star"], dtype=dt)
    np.testing.assert_array_equal(u @ S, np.array(g["u_lambda"]).reshape(1, 5))
    np.testing.assert_array_equal(S @ v, np.array(g["v_lambda"]).reshape(5, 1))


def test_vector_products(golden):
    g = golden["vector_product"]
    assert rb.array(g["u"]) @ rb.array(g["v"]) == g["result"]
    g = golden["vecmat"]
    np.testing.assert_array_equal(rb.array(g["u"]) @ rb.array(g["A"]), g["uA"])
    np.testing.assert_array_equal(np.array(g["u"]) @ rb.array(g["A"]), g["uA"])
    np.testing.assert_array_equal(rb.array(g["A"]) @ rb.array(g["v"]), g["Av"])


def test_matmul_random_vs_oracle(orc):
    rng = np.random.default_rng(2)
    A = synth.dense_pedigree(synth.PedigreeSpec(200, 8, 0.8, 3), 1)
    W = orc.warshall(A).W
    B = np.where(rng.random((200, 70)) < 0.1, rng.integers(2, 30, (200, 70)), 0).astype(np.int32)
    ref, ev = orc.matmul(W, B)
    np.testing.assert_array_equal(rb.array(W) @ rb.array(B), ref)


def test_compositions(golden):
    for key in ("vertex_compose_1", "vertex_compose_2", "vertex_compose_14"):
        g = golden[key]
        A = rb.array(g["A"], dtype=np.int32)
        out = rb.vertex_relational_composition(np.array(g["u"]).reshape(1, -1), A, np.array(g["v"]).reshape(-1, 1),
                                               g["color"])
        np.testing.assert_array_equal(out, g["out"])
    g = golden["vertex_cycle"]
    A_star = rb.array(g["A"]).transitive_closure().W
    with pytest.raises(ValueError) as e:
        A_star.vertex_relational_composition(rb.array(g["u"]), rb.array(g["v"]), g["color"])
    assert str(e.value) == g["message"]
    for key in ("edge_compose_simple", "edge_compose_15"):
        g = golden[key]
        out = rb.edge_relational_composition(rb.array(g["R"]), g["alpha"], g["beta"], g["rel"])
        np.testing.assert_array_equal(out, g["out"])


def test_edge_composition_equals_reclosure(orc):
    """Incremental edge insertion == closing the graph with the edge added."""
    A = synth.dense_pedigree(synth.PedigreeSpec(300, 10, 0.7, 3), 21)
    R = orc.warshall(A).W
    rng = np.random.default_rng(3)
    done = 0
    for _ in range(200):
        c, p = int(rng.integers(50, 300)), int(rng.integers(0, 50))
        if A[c, p] == 0 and R[p, c] == 0 and not A[c].tolist().count(2 if A[p, p] == -1 else 3):
            rel = 2 if A[p, p] == -1 else 3
            A2 = A.copy()
            A2[c, p] = rel
            exp = orc.warshall(A2).W
            got = rb.edge_relational_composition(rb.array(R), c, p, rel)
            np.testing.assert_array_equal(got, exp)
            done += 1
            if done == 3:
                break
    assert done > 0


@pytest.mark.parametrize("dt,n", [(np.int32, 1100), (np.int64, 700)])
def test_edge_compose_many_equals_reclosure(orc, dt, n):
    """A run of edge insertions on the device (no host round trips) equals
    closing the graph with every edge added; an edge that would close a cycle
    flags the run."""
    A = synth.dense_pedigree(synth.component_spec(n), 4, dtype=dt)
    et = elem_type(A.dtype)
    W, _ = dense.fw_closure(_lib.to_device(A), et)
    rng = np.random.default_rng(8)
    diag = np.diagonal(A)
    edges = []
    A2 = A.copy()
    for _ in range(16):
        a = int(rng.integers(n // 2, n))
        rel = int(rng.choice([2, 3]))
        cand = np.nonzero(diag[:a] == (-1 if rel == 2 else 1))[0]
        b = int(rng.choice(cand))
        edges.append((a, b, rel))
        A2[a, b] = rel
    got, st = dense.edge_compose_many(W.contiguous(), edges, et)
    assert not (st.flags & _lib.F_CYCLE) and st.events == 0
    if n <= 700:
        exp = orc.warshall(A2).W
    else:
        exp = dense.fw_closure(_lib.to_device(A2), et)[0].cpu().numpy()
    np.testing.assert_array_equal(got.cpu().numpy(), exp)
    # descendant -> ancestor is fine; ancestor -> descendant closes a cycle
    Wc = got.cpu().numpy()
    i, j = [int(x) for x in np.argwhere(np.tril(Wc, -1) > 0)[0]]
    _, st = dense.edge_compose_many(got, [edges[0], (j, i, 2)], et)
    assert st.flags & _lib.F_CYCLE


@pytest.mark.parametrize("dt,n", [(np.int32, 600), (np.int64, 400)])
def test_edge_compose_inplace_equals_out_of_place(dt, n):
    """The in-place composition (affected rows only) writes exactly what the
    reference-shaped out-of-place entry writes, edge after edge, including
    the composed row itself and rows that do not reach the edge's source."""
    import torch
    A = synth.dense_pedigree(synth.component_spec(n), 6, dtype=dt)
    et = elem_type(A.dtype)
    W, _ = dense.fw_closure(_lib.to_device(A), et)
    rng = np.random.default_rng(12)
    diag = np.diagonal(A)
    cur = W.contiguous()
    inpl = W.contiguous().clone()
    scratch = torch.empty(int(_lib.load().rbg_edge_compose_scratch_bytes(n)), dtype=torch.uint8, device=W.device)
    sfx = "i32" if dt == np.int32 else "i64"
    for _ in range(10):
        a = int(rng.integers(n // 2, n))
        rel = int(rng.choice([2, 3]))
        b = int(rng.choice(np.nonzero(diag[:a] == (-1 if rel == 2 else 1))[0]))
        st = dense.status()
        st.reset()
        out = torch.empty_like(cur)
        _lib.call(f"rbg_edge_compose_{sfx}", _lib.ptr(cur), n, a, b, rel, _lib.ptr(out), et.width, 0, st.p, _lib.stream_handle(None))
        s1 = st.read()
        st.reset()
        _lib.call(f"rbg_edge_compose_inplace_{sfx}", _lib.ptr(inpl), n, a, b, rel, _lib.ptr(scratch), et.width, 0,
                  st.p, _lib.stream_handle(None))
        s2 = st.read()
        assert (s1.flags, s1.events) == (s2.flags, s2.events)
        np.testing.assert_array_equal(inpl.cpu().numpy(), out.cpu().numpy())
        cur = out


@pytest.mark.parametrize("dt,n", [(np.int32, 33), (np.int8, 9), (np.int64, 65)])
def test_fw_overflow_chain(orc, dt, n):
    A = overflow_chain(n, dt)
    rep = orc.warshall(A).report
    assert rep["n_overflow"] > 0 and not rep["hang"] and rep["last_is_overflow"]
    with pytest.raises(OverflowError):
        rb.array(A).transitive_closure()
    W, out = _fw(A)
    assert out.mode == "general" and out.status.events != 0


# ---------------------------------------------------------------------------
# dense AVOS product, fast path
# ---------------------------------------------------------------------------
def _mm_call(name_sfx, A, B, fast):
    """C = A (x) B through one C-ABI entry (fast: rbg_matmul_fast_*, else the
    general rbg_matmul_*).  Returns (C, status)."""
    t = _lib.torch()
    M, K = A.shape
    N = B.shape[1]
    dA, dB = _lib.to_device(A), _lib.to_device(B)
    C = t.empty((M, N), dtype=dA.dtype, device=dA.device)
    st = dense.status()
    st.reset()
    sh = _lib.stream_handle(None)
    if fast:
        ws = t.empty(int(_lib.load().rbg_matmul_fast_workspace_bytes(M, K, N, dA.element_size())), dtype=t.uint8,
                     device=dA.device)
        _lib.call(f"rbg_matmul_fast_{name_sfx}", _lib.ptr(dA), K, _lib.ptr(dB), N, _lib.ptr(C), N, M, K, N,
                  _lib.ptr(ws), ws.numel(), st.p, sh)
    else:
        width = 32 if name_sfx == "i32" else 64
        _lib.call(f"rbg_matmul_{name_sfx}", _lib.ptr(dA), K, _lib.ptr(dB), N, _lib.ptr(C), N, M, K, N, width, 0,
                  st.p, sh)
    return C.cpu().numpy(), st.read()


@pytest.mark.parametrize("dt,n,shape", [(np.int32, 300, "square"), (np.int32, 300, "rect"), (np.int64, 200, "square"),
                                        (np.int64, 200, "rect"), (np.int32, 70, "square")])
def test_matmul_fast_vs_general_and_oracle(orc, dt, n, shape):
    """Closures of pedigrees (the fast domain): the fast entry equals the
    general kernel and the oracle's 'ij,jk' loop, for square and rectangular
    operands whose sizes are not multiples of the tile edges (diagonal terms
    of A and B fall inside and outside the product's range)."""
    A = synth.dense_pedigree(synth.component_spec(n, 0.8, 4), 41 + n, dtype=dt)
    W = orc.warshall(A).W
    if shape == "square":
        X, Y = W, W
    else:
        # leading blocks keep the diagonal on the diagonal (the fast domain)
        X, Y = W[: n - 37, : n - 11], W[: n - 11, : n - 5]
    sfx = "i32" if dt == np.int32 else "i64"
    st = dense.status()
    st.reset()
    for M in (X, Y):
        _lib.call(f"rbg_matmul_precheck_{sfx}", _lib.ptr(_lib.to_device(np.ascontiguousarray(M))), M.shape[1],
                  M.shape[0], M.shape[1], st.p, _lib.stream_handle(None))
    assert not st.read().flags & _lib.F_PRECOND_FAIL
    Cf, sf = _mm_call(sfx, np.ascontiguousarray(X), np.ascontiguousarray(Y), True)
    Cg, sg = _mm_call(sfx, np.ascontiguousarray(X), np.ascontiguousarray(Y), False)
    assert not sf.flags & _lib.F_NEED_GENERAL and sg.events == 0
    np.testing.assert_array_equal(Cf, Cg)
    ref, ev = orc.matmul(np.ascontiguousarray(X), np.ascontiguousarray(Y))
    np.testing.assert_array_equal(Cf, ref)


def test_matmul_fast_dispatch_and_fallbacks(orc):
    """rb.array @ on a closure large enough for the fast path equals the
    oracle; operands outside the domain (an off-diagonal 1) and products
    whose terms could overflow (long paternal chains) take the general
    kernel, with its events."""
    import redblackgraph_amd as rb
    A = synth.dense_pedigree(synth.component_spec(400, 0.8, 4), 5)
    W = orc.warshall(A).W
    ref, _ = orc.matmul(W, W)
    np.testing.assert_array_equal(rb.array(W) @ rb.array(W), ref)
    C, s = dense.matmul(_lib.to_device(W), _lib.to_device(W), elem_type(np.int32))
    np.testing.assert_array_equal(C.cpu().numpy(), ref)
    X = W.copy()
    X[300, 20] = 1  # outside the domain
    ref2, _ = orc.matmul(X, W)
    C2, _ = dense.matmul(_lib.to_device(X), _lib.to_device(W), elem_type(np.int32))
    np.testing.assert_array_equal(C2.cpu().numpy(), ref2)
    # a 31-generation paternal chain: its closure squared stays within int32
    # (W[i][k] (x) W[k][j] = 2^(i-j)), so the guard passes it and the result
    # is the oracle's; times its transpose the terms 2^(i-k) (x) 2^(j-k)
    # overflow, the guard sends it to the general kernel, which raises
    n = 256
    Ch = np.zeros((n, n), dtype=np.int32)
    np.fill_diagonal(Ch, -1)
    for i in range(1, 31):
        Ch[i, i - 1] = 2
    Wc = orc.warshall(Ch).W
    refc, evc = orc.matmul(Wc, Wc)
    assert evc == 0
    Cc, sc = _mm_call("i32", Wc, Wc, True)
    assert not sc.flags & (_lib.F_NEED_GENERAL | _lib.F_PRECOND_FAIL)
    np.testing.assert_array_equal(Cc, refc)
    WcT = np.ascontiguousarray(Wc.T)
    assert orc.matmul(Wc, WcT)[1] & 1  # the width overflow event
    Cf, sf = _mm_call("i32", Wc, WcT, True)
    assert sf.flags & _lib.F_NEED_GENERAL
    with pytest.raises(OverflowError):
        rb.array(Wc) @ rb.array(WcT)
