"""Tile-shape and schedule autotuner for the batched dense closure.

For each size bucket (padded order, a multiple of 128) the batched launcher
can run phase 3 with 128 x 128 tiles (8 x 8 cells per thread) or 64 x 64
tiles (4 x 4 cells per thread): the larger tile does fewer LDS reads per
relaxation, the smaller one gives a small component four times as many
workgroups per pivot block.  The tuner closes a batch of seeded synthetic
pedigrees of the bucket's order with each shape, times it with HIP events,
and keeps the faster one.

For a batch of lower triangular components it also picks the schedule:
the per-step launches or the persistent launch (BatchedClosure's
``persistent``).  Which one is faster depends on the batch's mix of orders
-- the per-step launches of many orders fill each other's tails, the
persistent launch wins on batches of large components -- so the tuner times
both on synthetic pedigrees of the batch's padded orders, keyed by their
histogram.

Results are cached as JSON, keyed by device name and bucket, in
$RBG_AUTOTUNE_CACHE (default ~/.cache/redblackgraph_amd/autotune.json); a
table shipped with the package (autotune_defaults.json) seeds the cache.
"""
from __future__ import annotations

import hashlib
import json
import os
import threading
from typing import Dict, Iterable, List

import numpy as np

from . import _lib

TILES = (128, 64)
_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULTS_PATH = os.path.join(_HERE, "autotune_defaults.json")
_lock = threading.Lock()


def cache_path() -> str:
    return os.environ.get("RBG_AUTOTUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "redblackgraph_amd", "autotune.json"))


def device_key() -> str:
    """Architecture and CU count (the marketing name is left out: some
    runtime configurations report it empty)."""
    t = _lib.torch()
    props = t.cuda.get_device_properties(_lib.device())
    arch = getattr(props, "gcnArchName", "") or ""
    return f"{arch.split(':')[0]}|{props.multi_processor_count}cu"


def _load(path: str) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path: str, table: dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs re-tuning


def _key(dev: str, npad: int) -> str:
    return f"{dev}|batched_fw_tile|{npad}"


def tuning_components(npad: int) -> int:
    """Components per timing batch: enough 128-tiles per pivot step for about
    sixteen rounds of workgroups over 256 CUs at three per CU, at most 4096
    components.  In a mixed batch the GPU is full while the large buckets
    run, so a shape is judged by its time at saturation, where the rounding
    of a small grid to whole rounds no longer favours the smaller tile."""
    tiles = (npad // 128) ** 2
    return int(max(1, min(4096, -(-16 * 3 * 256 // tiles))))


def measure_bucket(npad: int, reps: int = 2, seed: int = 7) -> Dict[int, float]:
    """Milliseconds per batched run of tuning_components(npad) components of
    order npad, for each tile shape."""
    from . import synth
    from .batched import BatchedClosure
    t = _lib.torch()
    m = tuning_components(npad)
    # a few distinct pedigrees, repeated: the timing does not depend on which
    distinct = [synth.component_edges(synth.component_spec(npad), seed * 1009 + i) for i in range(min(m, 8))]
    diag = [distinct[i % len(distinct)][0] for i in range(m)]
    edges = [distinct[i % len(distinct)][1] for i in range(m)]
    out = {}
    for tile in TILES:
        b = BatchedClosure([npad] * m, tiles={npad: tile})
        try:
            b.load_edges(diag, edges)
            b.run(graph=False)  # warm
            ev0, ev1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            ev0.record(b.stream)
            for _ in range(reps):
                b.run(graph=False)
            ev1.record(b.stream)
            ev1.synchronize()
            out[tile] = ev0.elapsed_time(ev1) / reps
        finally:
            b.close()
    return out


def tile_plan(npads: Iterable[int], measure: bool = True, retune: bool = False, dev: str | None = None) -> Dict[int, int]:
    """Tile edge per bucket for this device (or the device named ``dev``),
    measuring buckets not yet known when ``measure`` is set; unknown buckets
    default to 128."""
    npads: List[int] = sorted(set(int(p) for p in npads))
    dev = dev or device_key()
    with _lock:
        table = _load(DEFAULTS_PATH)
        table.update(_load(cache_path()))
        missing = [p for p in npads if retune or _key(dev, p) not in table]
        if missing and measure:
            for p in missing:
                ms = measure_bucket(p)
                best = min(ms, key=ms.get)
                table[_key(dev, p)] = {"tile": int(best), "ms": {str(k): round(v, 4) for k, v in ms.items()},
                                       "components": tuning_components(p)}
            _save(cache_path(), table)
    plan = {}
    for p in npads:
        e = table.get(_key(dev, p))
        plan[p] = int(e["tile"]) if e else 128
    return plan


SCHEDULES = ("steps", "persistent")


def _schedule_key(dev: str, npads: Iterable[int]) -> str:
    u, c = np.unique(np.asarray(list(npads), dtype=np.int64), return_counts=True)
    digest = hashlib.sha1(",".join(f"{a}:{b}" for a, b in zip(u, c)).encode()).hexdigest()[:16]
    return f"{dev}|batched_fw_schedule|{digest}"


def measure_schedule(npads: Iterable[int], tiles: Dict[int, int], reps: int = 2, seed: int = 7) -> Dict[str, float]:
    """Milliseconds per eager run of a batch of lower triangular synthetic
    pedigrees with these padded orders, per schedule."""
    from . import synth
    from .batched import BatchedClosure
    t = _lib.torch()
    npads = [int(p) for p in npads]
    pool = {p: [synth.component_edges(synth.component_spec(p), seed * 1009 + i) for i in range(4)]
            for p in sorted(set(npads))}
    diag = [pool[p][i % 4][0] for i, p in enumerate(npads)]
    edges = [pool[p][i % 4][1] for i, p in enumerate(npads)]
    out = {}
    for name in SCHEDULES:
        b = BatchedClosure(npads, tiles=tiles, tri=1, persistent=name == "persistent")
        try:
            b.load_edges(diag, edges)
            b.run(graph=False)  # warm
            ev0, ev1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
            ev0.record(b.stream)
            for _ in range(reps):
                b.run(graph=False)
            ev1.record(b.stream)
            ev1.synchronize()
            out[name] = ev0.elapsed_time(ev1) / reps
        finally:
            b.close()
            del b
    return out


def schedule_plan(npads: Iterable[int], tiles: Dict[int, int], measure: bool = True, retune: bool = False,
                  dev: str | None = None) -> str:
    """"steps" or "persistent" for a batch of lower triangular components of
    these padded orders on this device (or the device named ``dev``),
    measuring an unknown mix when ``measure`` is set; "steps" when unknown."""
    npads = [int(p) for p in npads]
    dev = dev or device_key()
    key = _schedule_key(dev, npads)
    with _lock:
        table = _load(DEFAULTS_PATH)
        table.update(_load(cache_path()))
        if (retune or key not in table) and measure:
            ms = measure_schedule(npads, tiles)
            table[key] = {"schedule": min(ms, key=ms.get), "ms": {k: round(v, 3) for k, v in ms.items()},
                          "components": len(npads)}
            _save(cache_path(), table)
        e = table.get(key)
    return e["schedule"] if e else "steps"
