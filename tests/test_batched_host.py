"""Host-side logic of the batched closure: pool layout and the autotuner's
cache (no GPU calls)."""
import json

import numpy as np

from redblackgraph_amd import autotune
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
