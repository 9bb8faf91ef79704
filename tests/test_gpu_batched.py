"""GPU parity of the batched dense closure: every component equals the CPU
oracle's naive loop (W and diameter) and the single-matrix GPU path, for both
tile shapes, eagerly and through hipGraph replays; the persistent launch of
lower triangular components equals the per-step launches; components outside
the fast domain fall back to the general kernels with the single-matrix
errors."""
import numpy as np
import pytest

from redblackgraph_amd import _lib, dense, synth
from redblackgraph_amd._dtypes import elem_type, msb
from redblackgraph_amd.batched import BatchedClosure, transitive_closure_batch
from helpers import overflow_chain

pytestmark = pytest.mark.gpu


def _components(count, lo, hi, seed):
    sizes = synth.batched_sizes(count, lo, hi, seed)
    return [synth.dense_pedigree(synth.component_spec(int(s)), synth.component_seed(seed, i))
            for i, s in enumerate(sizes)]


@pytest.mark.parametrize("tile", [128, 64])
def test_batched_vs_oracle(orc, tile):
    mats = _components(40, 16, 400, 1)
    b = BatchedClosure([A.shape[0] for A in mats], tiles={p: tile for p in range(128, 4097, 128)})
    try:
        b.load_dense(mats)
        b.run(graph=False)
        outs = b.outcomes()
        for i, A in enumerate(mats):
            ref = orc.warshall(A)
            assert outs[i].mode == "fast"
            np.testing.assert_array_equal(b.result(i).cpu().numpy(), ref.W, err_msg=f"component {i}")
            assert outs[i].diameter == ref.diameter, i
            assert outs[i].status.cycle_count == 0
    finally:
        b.close()


def test_batched_equals_single_matrix_path_with_graph_replays():
    """A few hundred components up to 2048: the batched result equals
    dense.fw_closure per component; graph replays, with eager work and
    synchronisations between them, reproduce the eager run every time."""
    t = _lib.torch()
    sizes, diag, edges = synth.batched_workload(count=160, lo=16, hi=2048, seed=3)
    b = BatchedClosure(sizes)
    try:
        b.load_edges(diag, edges)
        b.run(graph=False)
        b.stream.synchronize()  # run() is asynchronous on the batch's stream
        eager = b.pool.clone()
        st0 = [(s.diameter_max, s.cycle_count, s.flags) for s in b.statuses()]
        et = elem_type(np.int32)
        for i in range(0, len(sizes), 7):
            n = int(sizes[i])
            A = t.zeros((n, n), dtype=t.int32, device=b.pool.device)
            r, c, v = edges[i]
            A[t.from_numpy(r).to(A.device), t.from_numpy(c).to(A.device)] = t.from_numpy(v.astype(np.int32)).to(A.device)
            A.diagonal().copy_(t.from_numpy(diag[i]).to(A.device))
            W, out = dense.fw_closure(A, et)
            assert bool(t.equal(W, b.result(i))), i
            assert out.status.diameter_max == st0[i][0]
        for rep in range(3):
            b.load_edges(diag, edges)
            b.run(graph=True)
            b.stream.synchronize()
            assert bool(t.equal(b.pool, eager)), f"replay {rep}"
            assert [(s.diameter_max, s.cycle_count, s.flags) for s in b.statuses()] == st0
            _ = (t.arange(1000, device=b.pool.device) * 3).sum().item()  # unrelated eager work
    finally:
        b.close()


def test_batched_fallbacks_and_errors(orc):
    """Components the fast kernels cannot close (an off-diagonal 1, a cycle, a
    value chain that overflows int32) come back through the general kernels
    with the single-matrix path's results and errors."""
    good = _components(6, 16, 300, 5)
    off1 = good[1].copy()
    off1[5, 2] = 1
    cyc = good[2].copy()
    i, j = np.nonzero(np.tril(cyc, -1))
    cyc[j[0], i[0]] = 2
    chain = overflow_chain(33, np.int32)
    mats = [good[0], off1, cyc, chain, good[4], good[5]]
    b = BatchedClosure([A.shape[0] for A in mats])
    try:
        b.load_dense(mats)
        b.run(graph=False)
        sts = b.statuses()
        assert sts[1].flags & _lib.F_PRECOND_FAIL
        assert sts[2].cycle_count > 0 or sts[2].flags & _lib.F_NEED_GENERAL
        assert sts[3].flags & _lib.F_NEED_GENERAL
        outs = b.outcomes(lambda q: _lib.to_device(mats[q]))
        et = elem_type(np.int32)
        for q, A in enumerate(mats):
            W, single = dense.fw_closure(_lib.to_device(A), et)
            np.testing.assert_array_equal(b.result(q).cpu().numpy(), W.cpu().numpy(), err_msg=f"component {q}")
            assert outs[q].mode == single.mode
            assert (outs[q].status.cycle_count, outs[q].status.cycle_last_key, outs[q].status.events) == \
                (single.status.cycle_count, single.status.cycle_last_key, single.status.events)
            if q not in (2, 3):
                ref = orc.warshall(A)
                np.testing.assert_array_equal(W.cpu().numpy(), ref.W)
                assert outs[q].diameter == ref.diameter
        assert [o.mode for o in outs] == ["fast", "general", "general", "general", "fast", "fast"]
    finally:
        b.close()
    with pytest.raises(ValueError, match="Error: cycle detected!"):
        transitive_closure_batch(mats)
    with pytest.raises(OverflowError):
        transitive_closure_batch([good[0], chain])
    res = transitive_closure_batch([good[0], off1])
    np.testing.assert_array_equal(res[1].W, orc.warshall(off1).W)
    assert res[0].diameter == orc.warshall(good[0]).diameter


@pytest.mark.parametrize("tile,reverse,persistent", [(128, False, False), (64, False, False), (128, True, False),
                                                     (128, False, True), (64, False, True)])
def test_batched_triangular_vs_oracle(orc, tile, reverse, persistent):
    """Components declared triangular take the triangular schedule: lower
    (parents first) and upper (numbered backwards) pedigrees equal the
    oracle's naive loop, W and trajectory diameter, by per-step launches and
    by the persistent launch."""
    mats = _components(40, 16, 600, 7)
    if reverse:
        mats = [synth.reverse_vertices(A) for A in mats]
    b = BatchedClosure([A.shape[0] for A in mats], tiles={p: tile for p in range(128, 4097, 128)},
                       tri=2 if reverse else 1, persistent=persistent)
    assert b.info()["persistent_components"] == (len(mats) if persistent else 0)
    try:
        b.load_dense(mats)
        b.run(graph=False)
        outs = b.outcomes()
        for i, A in enumerate(mats):
            ref = orc.warshall(A)
            assert outs[i].mode == "fast", i
            np.testing.assert_array_equal(b.result(i).cpu().numpy(), ref.W, err_msg=f"component {i}")
            assert outs[i].diameter == ref.diameter, i
    finally:
        b.close()


@pytest.mark.parametrize("persistent", [False, True])
def test_batched_triangular_large_components_and_replays(persistent):
    """Components above 2048 vertices (where most of the batched work is)
    through the triangular schedule, per step and persistent: equal to the
    single-matrix path (whose 128-tile kernel the oracle digests pin),
    eagerly and in graph replays; a component declared triangular that is
    not is caught and closed again."""
    t = _lib.torch()
    sizes, diag, edges = synth.batched_workload(count=24, lo=1500, hi=4096, seed=9)
    b = BatchedClosure(sizes, tri=1, persistent=persistent)
    try:
        b.load_edges(diag, edges)
        b.run(graph=False)
        b.stream.synchronize()
        eager = b.pool.clone()
        outs = b.outcomes()
        et = elem_type(np.int32)
        for i in range(len(sizes)):
            n = int(sizes[i])
            A = t.zeros((n, n), dtype=t.int32, device=b.pool.device)
            r, c, v = edges[i]
            A[t.from_numpy(r).to(A.device), t.from_numpy(c).to(A.device)] = t.from_numpy(v.astype(np.int32)).to(A.device)
            A.diagonal().copy_(t.from_numpy(diag[i]).to(A.device))
            W, out = dense.fw_closure(A, et, schedule="full")
            assert outs[i].mode == "fast"
            assert bool(t.equal(W, b.result(i))), i
            assert out.status.diameter_max == outs[i].status.diameter_max, i
        for rep in range(2):
            b.load_edges(diag, edges)
            b.run(graph=True)
            b.stream.synchronize()
            assert bool(t.equal(b.pool, eager)), f"replay {rep}"
    finally:
        b.close()
    # a shuffled (non-triangular) component declared lower triangular
    mats = _components(3, 100, 300, 11)
    bad = synth.shuffle_vertices(mats[1], 3)
    mats = [mats[0], bad, mats[2]]
    res = transitive_closure_batch(mats, tri=1)
    W, out = dense.fw_closure(_lib.to_device(bad), elem_type(np.int32))
    np.testing.assert_array_equal(res[1].W, W.cpu().numpy())
    assert res[1].diameter == out.diameter


def test_batched_persistent_equals_per_step_launches(orc):
    """Lower triangular components in the persistent launch (128- and
    64-tiles in one grid), beside full-schedule components in per-step
    launches of the same batch: every matrix and status equals the per-step
    schedule's, eagerly and in graph replays, and the oracle's on a sample."""
    t = _lib.torch()
    mats = _components(48, 16, 1300, 21)
    tri = [0 if i % 3 == 0 else 1 for i in range(len(mats))]
    tiles = {p: (64 if (p // 128) % 2 else 128) for p in range(128, 4097, 128)}
    runs = {}
    for persistent in (True, False, 512):
        b = BatchedClosure([A.shape[0] for A in mats], tiles=tiles, tri=tri, persistent=persistent)
        try:
            info = b.info()
            npad = [dense.padded_order(A.shape[0]) for A in mats]
            want = {True: sum(tri), False: 0, 512: sum(t_ and p >= 512 for t_, p in zip(tri, npad))}[persistent]
            assert info["persistent_components"] == want
            b.load_dense(mats)
            b.run(graph=False)
            b.stream.synchronize()
            eager = b.pool.clone()
            sts = [(s.flags, s.diameter_max, s.cycle_count, s.events) for s in b.statuses()]
            if persistent:
                for rep in range(3):
                    b.load_dense(mats)
                    b.run(graph=True)
                    b.stream.synchronize()
                    assert bool(t.equal(b.pool, eager)), f"replay {rep}"
                    assert [(s.flags, s.diameter_max, s.cycle_count, s.events) for s in b.statuses()] == sts
            runs[persistent] = (eager, sts, b.layout)
        finally:
            b.close()
    (pool_p, st_p, lay), (pool_s, st_s, _) = runs[True], runs[False]
    assert st_p == st_s == runs[512][1]
    assert bool(t.equal(runs[512][0], pool_s))
    assert all(f == 0 for f, _, _, _ in st_p)
    assert bool(t.equal(pool_p, pool_s))
    for i in (1, 2, 4, 5, 7):
        n, o, p = int(lay.sizes[i]), int(lay.offset[i]), int(lay.npad[i])
        ref = orc.warshall(mats[i])
        np.testing.assert_array_equal(pool_p[o:o + p * p].view(p, p)[:n, :n].cpu().numpy(), ref.W)
        assert msb(int(st_p[i][1])) == ref.diameter, i


def test_batched_persistent_wait_limit_reruns():
    """Waits of the persistent batched launch that run out (a one-tick limit
    set from the host) flag their components stalled; outcomes() closes those
    again through the single-matrix path, and every component equals an
    unforced run."""
    t = _lib.torch()
    L = _lib.load()
    mats = _components(24, 200, 1500, 33)
    b = BatchedClosure([A.shape[0] for A in mats], tri=1, persistent=True)
    try:
        assert b.info()["persistent_components"] == len(mats)
        b.load_dense(mats)
        b.run(graph=False)
        outs0 = b.outcomes()
        ref = [b.result(i).clone() for i in range(len(mats))]
        prev = L.rbg_fw_flow_wait_limit(1)
        try:
            b.load_dense(mats)
            b.run(graph=False)
            stalled = [i for i, s in enumerate(b.statuses()) if s.flags & _lib.F_STALLED]
            outs1 = b.outcomes(lambda i: _lib.to_device(mats[i]))
        finally:
            L.rbg_fw_flow_wait_limit(prev)
        assert stalled
        for i in range(len(mats)):
            assert bool(t.equal(b.result(i), ref[i])), i
            assert outs1[i].diameter == outs0[i].diameter, i
    finally:
        b.close()


def test_batched_schedule_autotuned(tmp_path, monkeypatch):
    """persistent="auto": the autotuner times both schedules on synthetic
    pedigrees of the batch's padded orders, caches the faster one, and the
    batch closed by it equals the per-step launches' closure."""
    from redblackgraph_amd import autotune
    t = _lib.torch()
    monkeypatch.setenv("RBG_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    mats = _components(20, 100, 1300, 41)
    b = BatchedClosure([A.shape[0] for A in mats], tri=1, persistent="auto")
    s = BatchedClosure([A.shape[0] for A in mats], tri=1)
    try:
        assert b.schedule in autotune.SCHEDULES
        assert b.info()["persistent_components"] == (len(mats) if b.schedule == "persistent" else 0)
        key = autotune._schedule_key(autotune.device_key(), b.layout.npad)
        import json
        e = json.load(open(tmp_path / "tune.json"))[key]
        assert e["schedule"] == b.schedule and set(e["ms"]) == set(autotune.SCHEDULES)
        for x in (b, s):
            x.load_dense(mats)
            x.run(graph=False)
            x.stream.synchronize()
        assert bool(t.equal(b.pool, s.pool))
    finally:
        b.close()
        s.close()
