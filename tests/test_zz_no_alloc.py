"""The allocation-free dense closure (rbg_fw_dense_noalloc_*) against the
entry point it is a variant of (rbg_fw_dense_*, caller workspace): the same
closed matrix and the same status block, through the C ABI, on two calls in
a row with the second one smaller (the device array keeps the first call's
contents), in fast and general mode, int32 and int64; sizes beyond the
array are refused before anything runs."""
import numpy as np
import pytest

from redblackgraph_amd import _lib, dense, synth

pytestmark = pytest.mark.gpu


def _bind():
    return _lib.load()


def _status_tuple(s):
    return (s.events, s.flags, s.cycle_count, s.cycle_first_key, s.cycle_last_key, s.ov_first_key,
            s.ov_last_key, s.cycle_first_a, s.cycle_first_b, s.cycle_last_a, s.cycle_last_b,
            s.ov_last_a, s.ov_last_b, s.diameter_max)


def _close_both(A, mode):
    """Closes A with the workspace entry and the allocation-free entry."""
    L = _bind()
    t = _lib.torch()
    n = A.shape[0]
    assert n % 128 == 0
    sfx = "i32" if A.dtype == np.int32 else "i64"
    width = 32 if A.dtype == np.int32 else 64
    elem = A.dtype.itemsize
    W1 = _lib.to_device(A)
    W2 = W1.clone()
    ws = t.empty(int(L.rbg_fw_workspace_bytes(n, elem)), dtype=t.uint8, device=W1.device)
    sh = _lib.stream_handle(None)
    st1, st2 = _lib.Status(), _lib.Status()
    st1.reset()
    _lib.call(f"rbg_fw_dense_{sfx}", _lib.ptr(W1), n, n, width, 0, mode, _lib.ptr(ws), ws.numel(), st1.p, sh)
    s1 = st1.read()
    st2.reset()
    rc = getattr(L, f"rbg_fw_dense_noalloc_{sfx}")(_lib.ptr(W2), n, n, width, 0, mode, st2.p, sh)
    assert rc == 0
    s2 = st2.read()
    return W1.cpu().numpy(), s1, W2.cpu().numpy(), s2


@pytest.mark.parametrize("dt,sizes", [(np.int32, (1536, 512)), (np.int64, (1024, 384))])
def test_noalloc_fast_matches_workspace_entry(orc, dt, sizes):
    for call, n in enumerate(sizes):
        spec = synth.PedigreeSpec(n, 12, 0.75, 4)
        A = synth.dense_pedigree(spec, 40 + call, dtype=dt)
        print(f"noalloc fast: {np.dtype(dt).name} n={n} call {call}", flush=True)
        W1, s1, W2, s2 = _close_both(A, 0)
        np.testing.assert_array_equal(W2, W1)
        assert _status_tuple(s2) == _status_tuple(s1)
        assert not (s1.flags & _lib.F_NEED_GENERAL)
        if n <= 512:
            ref = orc.warshall(A)
            np.testing.assert_array_equal(W2, ref.W)
            assert dense.msb(s2.diameter_max) == ref.diameter


def test_noalloc_general_matches_workspace_entry():
    """General mode (full arithmetic, every event keyed) on a cyclic pedigree,
    then a smaller acyclic one."""
    # back edges added to a pedigree (the case of test_fw_cyclic_pedigree_errors),
    # padded with zero rows and columns to 384
    A0 = synth.dense_pedigree(synth.PedigreeSpec(260, 10, 0.8, 3), 5)
    rng = np.random.default_rng(5)
    for _ in range(6):
        i, j = sorted(rng.integers(0, 260, 2))
        if i != j:
            A0[i, j] = 2 if A0[j, j] == -1 else 3
    A = np.zeros((384, 384), dtype=np.int32)
    A[:260, :260] = A0
    print("noalloc general: int32 n=384 cyclic", flush=True)
    W1, s1, W2, s2 = _close_both(A, 1)
    assert s1.cycle_count > 0
    np.testing.assert_array_equal(W2, W1)
    assert _status_tuple(s2) == _status_tuple(s1)
    B = synth.dense_pedigree(synth.PedigreeSpec(256, 8, 0.8, 3), 6)
    print("noalloc general: int32 n=256", flush=True)
    W1, s1, W2, s2 = _close_both(B, 1)
    np.testing.assert_array_equal(W2, W1)
    assert _status_tuple(s2) == _status_tuple(s1)


def test_noalloc_refuses_oversize():
    L = _bind()
    t = _lib.torch()
    cap = int(L.rbg_fw_noalloc_max_bytes())
    # the smallest padded order whose workspace exceeds the array
    n = 128
    while int(L.rbg_fw_workspace_bytes(n, 4)) <= cap:
        n += 128
    # a matrix of the full order, so that even a wrongly accepted call stays in bounds
    W = t.zeros((n, n), dtype=t.int32, device=_lib.device())
    st = _lib.Status()
    st.reset()
    rc = L.rbg_fw_dense_noalloc_i32(_lib.ptr(W), n, n, 32, 0, 0, st.p, _lib.stream_handle(None))
    assert rc == -2  # RBG_E_WORKSPACE, before anything runs
    assert int(L.rbg_fw_workspace_bytes(n - 128, 4)) <= cap  # n - 128 still fits
    assert n - 128 >= 16384
