"""Host-side logic of the batched closure: pool layout, the persistent
launch's tickets and the autotuner's cache (no GPU calls)."""
import ctypes
import json

import numpy as np
import pytest

from redblackgraph_amd import _lib, autotune
from redblackgraph_amd.batched import BatchLayout


def test_layout_offsets_and_padding():
    sizes = [16, 129, 4096, 300, 128]
    L = BatchLayout.build(sizes)
    assert list(L.npad) == [128, 256, 4096, 384, 128]
    ends = L.offset + L.npad ** 2
    assert L.offset[0] == 0 and np.all(L.offset[1:] == ends[:-1]) and L.pool_elems == ends[-1]
    assert np.all(L.offset % (128 * 128) == 0)
    assert np.all(L.snap_offset % 256 == 0)
    assert np.all(np.diff(L.snap_offset) >= 4 * 32 * L.npad[:-1] * 4)
    assert np.all(np.diff(L.snap_offset) >= L.snap_size[:-1])
    P = BatchLayout.build(sizes, [True, False, True, True, False])
    assert np.all(np.diff(P.snap_offset) >= P.snap_size[:-1]) and P.snap_offset[-1] + P.snap_size[-1] <= P.snap_bytes
    lib = _lib.load()
    for i, p in enumerate(P.npad):
        q = lib.rbg_batch_tri_snap_bytes(int(p)) if i in (0, 2, 3) else lib.rbg_batch_snap_bytes(int(p))
        assert P.snap_size[i] == q
        assert lib.rbg_batch_tri_snap_bytes(int(p)) > lib.rbg_batch_snap_bytes(int(p))


def _tickets(npad, edge):
    lib = _lib.load()
    a = np.asarray(npad, dtype=np.int32)
    e = np.asarray(edge, dtype=np.int32)
    vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    total = lib.rbg_batch_flow_tickets(vp(a), vp(e), len(a), None, 0)
    out = np.zeros(2 * total, dtype=np.int64)
    assert lib.rbg_batch_flow_tickets(vp(a), vp(e), len(a), vp(out), out.size) == total
    return out.reshape(total, 2)


@pytest.mark.parametrize("npad,edge", [
    ([4096, 2048, 1024, 1024, 512, 384, 128, 128, 128], [128] * 9),
    ([1280, 1280, 1152, 640, 256, 256, 128], [128, 64, 64, 128, 64, 64, 128]),
    ([128], [128]),
])
def test_persistent_batch_tickets(npad, edge):
    """Every ticket of the persistent batched launch names one work item of
    one component; each component's items all appear, once, in its own
    ticket order (so every wait, which names an item of the same component
    with a smaller own ticket, names a smaller ticket of the launch); and the
    components of one order and tile edge run in lockstep, largest first."""
    lib = _lib.load()
    tk = _tickets(npad, edge)
    comp, local = tk[:, 0], tk[:, 1]
    for c, (p, e) in enumerate(zip(npad, edge)):
        own = local[comp == c]
        total = lib.rbg_fw_flow_schedule(p, 4, e, None, 0)
        np.testing.assert_array_equal(own, np.arange(total))
    first = [np.flatnonzero(comp == c)[0] for c in range(len(npad))]
    assert first == sorted(first) and first[0] == 0
    n0 = [lib.rbg_fw_flow_schedule(p, 4, e, None, 0) for p, e in zip(npad, edge)]
    assert len(tk) == sum(n0)
    # lockstep: within a run of equal (order, edge) the components take turns
    x = 0
    while x < len(tk):
        c0 = comp[x]
        run = [c for c in range(c0, len(npad)) if (npad[c], edge[c]) == (npad[c0], edge[c0])]
        m = len(run)
        blk = comp[x:x + m * n0[c0]].reshape(n0[c0], m)
        np.testing.assert_array_equal(blk, np.tile(np.asarray(run), (n0[c0], 1)))
        x += m * n0[c0]


def test_persistent_batch_tickets_reject_bad_order():
    lib = _lib.load()
    a = np.asarray([128, 256], dtype=np.int32)
    e = np.asarray([128, 128], dtype=np.int32)
    vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    assert lib.rbg_batch_flow_tickets(vp(a), vp(e), 2, None, 0) < 0
    e = np.asarray([128, 96], dtype=np.int32)
    a = np.asarray([256, 128], dtype=np.int32)
    assert lib.rbg_batch_flow_tickets(vp(a), vp(e), 2, None, 0) < 0


def test_tuning_batch_sizes():
    assert autotune.tuning_components(128) == 4096
    assert autotune.tuning_components(4096) == 12
    assert autotune.tuning_components(1024) == 192


def test_tile_plan_reads_cache(tmp_path, monkeypatch):
    cache = tmp_path / "tune.json"
    dev = "TestGPU|gfx950|256cu"
    cache.write_text(json.dumps({f"{dev}|batched_fw_tile|256": {"tile": 64, "ms": {}}}))
    monkeypatch.setenv("RBG_AUTOTUNE_CACHE", str(cache))
    plan = autotune.tile_plan([256, 512, 256], measure=False, dev=dev)
    assert plan == {256: 64, 512: 128}


def test_schedule_plan_reads_cache(tmp_path, monkeypatch):
    """The schedule choice is keyed by the histogram of padded orders (the
    order of the components does not matter) and defaults to the per-step
    launches when the mix is unknown and not measured."""
    dev = "TestGPU|gfx950|256cu"
    npads = [128, 4096, 128, 2048]
    key = autotune._schedule_key(dev, npads)
    assert key == autotune._schedule_key(dev, sorted(npads)) != autotune._schedule_key(dev, npads[:3])
    cache = tmp_path / "tune.json"
    cache.write_text(json.dumps({key: {"schedule": "persistent", "ms": {}}}))
    monkeypatch.setenv("RBG_AUTOTUNE_CACHE", str(cache))
    assert autotune.schedule_plan(npads, {}, measure=False, dev=dev) == "persistent"
    assert autotune.schedule_plan([256, 256], {}, measure=False, dev=dev) == "steps"


def test_shipped_tile_table(tmp_path, monkeypatch):
    """The table shipped for MI355X (gfx950, 256 CUs) covers every padded
    order up to 4096 and seeds the plan without measuring."""
    table = json.load(open(autotune.DEFAULTS_PATH))
    dev = "gfx950|256cu"
    for npad in range(128, 4097, 128):
        e = table[f"{dev}|batched_fw_tile|{npad}"]
        assert e["tile"] in autotune.TILES
        assert e["components"] == autotune.tuning_components(npad)
    monkeypatch.setenv("RBG_AUTOTUNE_CACHE", str(tmp_path / "empty.json"))
    plan = autotune.tile_plan(range(128, 4097, 128), measure=False, dev=dev)
    assert set(plan.values()) <= set(autotune.TILES) and len(plan) == 32
    # the bench workload's schedule ships too, so a fresh install measures nothing
    from redblackgraph_amd import synth
    from redblackgraph_amd.dense import padded_order
    npads = [padded_order(int(n)) for n in synth.batched_workload(count=4096, seed=0)[0]]
    e = table[autotune._schedule_key(dev, npads)]
    assert e["schedule"] in autotune.SCHEDULES and e["schedule"] == min(e["ms"], key=e["ms"].get)


def test_serving_warmup_trace_shapes():
    """The serving warm-up covers the largest dense and sparse shapes, each
    closed and then given the maximum number of acyclic parent edges."""
    import scipy.sparse as sp
    from redblackgraph_amd.serving import DENSE_LIMIT, warmup_trace
    reqs = warmup_trace(hi=5000, max_edges=8, seed=3)
    assert [r.kind for r in reqs] == ["close", "insert", "close", "insert"]
    dense_g, sparse_g = reqs[0].graph, reqs[2].graph
    assert isinstance(dense_g, np.ndarray) and dense_g.shape == (DENSE_LIMIT - 1, DENSE_LIMIT - 1)
    assert sp.issparse(sparse_g) and sparse_g.shape == (5000, 5000)
    for r in (reqs[1], reqs[3]):
        assert 1 <= len(r.edges) <= 8
        assert all(b < a and rel in (2, 3) for a, b, rel in r.edges)
