"""Batched dense closure: thousands of independent components closed by one
launch schedule (recordable as one hipGraph) instead of one closure per graph.

The reference closes components one at a time
(sparse/csgraph/transitive_closure.py:85-195, component_wise_closure, each
component through the dense closure of core/redblack.py:49-50 /
warshall.c.in:50-107).  Here every component is padded to a multiple of 128
and laid out contiguously in one int32 pool in HBM; the C ABI
(rbg_batch_create / run / capture / replay) walks all components' pivot
blocks together, largest components first, with the phase-3 tile shape picked
per size bucket (see autotune.py).  Lower triangular components (pedigrees
numbered parents first) are instead closed together by one persistent launch:
every component's work items in one grid, each waiting only on its own
component's counters.

Each component's result equals the single-matrix closure
(dense.fw_closure): W, diameter, and the error behaviour.  Components whose
input is outside the fast kernels' domain, or that need general arithmetic or
contain a cycle, are closed again from their input by the single-matrix path.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _lib
from ._dtypes import elem_type, msb
from .dense import FwOutcome, fw_closure, padded_order

_ALIGN = 256


def _round(x: int, a: int = _ALIGN) -> int:
    return (x + a - 1) // a * a


@dataclass
class BatchLayout:
    """Where each component lives: pool offsets (elements), padded orders,
    snapshot offsets and sizes (bytes)."""
    sizes: np.ndarray
    npad: np.ndarray
    offset: np.ndarray
    snap_offset: np.ndarray
    snap_size: np.ndarray
    pool_elems: int
    snap_bytes: int

    @classmethod
    def build(cls, sizes: Sequence[int], persistent=None) -> "BatchLayout":
        """``persistent``: per component, whether its workspace should hold
        the persistent closure's (rbg_batch_tri_snap_bytes)."""
        L = _lib.load()
        sizes = np.asarray(sizes, dtype=np.int64)
        npad = np.array([padded_order(int(n)) for n in sizes], dtype=np.int64)
        pers = np.zeros(len(sizes), dtype=bool) if persistent is None else np.asarray(persistent, dtype=bool)
        off = np.zeros(len(sizes), dtype=np.int64)
        soff = np.zeros(len(sizes), dtype=np.int64)
        ssz = np.zeros(len(sizes), dtype=np.int64)
        pos = spos = 0
        snap_cache: Dict[tuple, int] = {}
        for i, p in enumerate(npad):
            off[i] = pos
            pos += int(p) * int(p)  # multiples of 128 x 128: every component stays 64 KiB aligned
            soff[i] = spos
            key = (int(p), bool(pers[i]))
            b = snap_cache.get(key)
            if b is None:
                q = L.rbg_batch_tri_snap_bytes if key[1] else L.rbg_batch_snap_bytes
                b = snap_cache[key] = int(q(int(p)))
            ssz[i] = b
            spos += _round(b)
        return cls(sizes, npad, off, soff, ssz, pos, spos)


class BatchedClosure:
    """Device state for closing a fixed list of component sizes repeatedly.

    ``tiles`` maps padded order -> phase-3 tile edge (128 or 64); by default
    the autotuner's choice for this device.  ``tri`` declares the components'
    triangularity (0: none, the full schedule; 1 / 2: every off-diagonal
    entry below / above the diagonal, the triangular schedule), one value for
    all or one per component; a component that is not triangular as declared
    is found by the device precheck and closed again by the single-matrix
    path, like any component outside the fast domain.  ``persistent``: lower
    triangular components (tri 1) are closed by one persistent launch
    instead of one panel and tile launches per pivot step (the same results);
    True for all of them, a padded order from which on (the others keep
    the per-step launches), or "auto": the autotuner's choice for this mix
    of orders (autotune.schedule_plan).  On MI355X the 4096-component
    workload takes 396 ms all persistent against 364 ms per step (every work
    item pays a few us of decode, polling and publication, and the per-step
    launches of a mixed batch fill each other's tails); batches of large
    components run 20-28% faster persistent.
    """

    def __init__(self, sizes: Sequence[int], tiles: Optional[Dict[int, int]] = None, stream=None, tri=0,
                 persistent=False):
        t = _lib.torch()
        dev = _lib.device()
        count = len(sizes)
        tri = np.broadcast_to(np.asarray(tri, dtype=np.int32), (count,))
        if np.any((tri < 0) | (tri > 2)):
            raise ValueError("tri must be 0, 1 or 2")
        self.tri = tri.copy()
        npad = np.array([padded_order(int(n)) for n in sizes], dtype=np.int64)
        if tiles is None:
            from .autotune import tile_plan
            tiles = tile_plan(sorted(set(int(p) for p in npad)))
        self.tiles = {int(k): int(v) for k, v in tiles.items()}
        if isinstance(persistent, str):
            if persistent != "auto":
                raise ValueError(f"unknown schedule {persistent!r}")
            from .autotune import schedule_plan
            lower = npad[self.tri == 1]
            persistent = bool(lower.size) and schedule_plan(lower, self.tiles) == "persistent"
        if isinstance(persistent, (bool, np.bool_)):
            pers = np.full(count, bool(persistent))
            self.schedule = "persistent" if persistent else "steps"
        else:
            pers = npad >= int(persistent)
            self.schedule = f"persistent from order {int(persistent)}"
        # triangular components get the persistent closure's workspace either
        # way: the per-step launches use its diagonal snapshots
        self.layout = BatchLayout.build(sizes, self.tri != 0)
        self.persistent_mask = (self.tri == 1) & pers
        self.count = len(self.layout.sizes)
        self.stream = stream if stream is not None else t.cuda.Stream()
        self.pool = t.zeros(self.layout.pool_elems, dtype=t.int32, device=dev)
        self.snap = t.empty(max(self.layout.snap_bytes, _ALIGN), dtype=t.uint8, device=dev)
        self.status = t.zeros(self.count * _lib.STATUS_BYTES, dtype=t.uint8, device=dev)
        self.status_host = t.empty(self.count * _lib.STATUS_BYTES, dtype=t.uint8, pin_memory=True)
        descs = (_lib.RbgBatchComp * self.count)()
        pb, sb, stb = self.pool.data_ptr(), self.snap.data_ptr(), self.status.data_ptr()
        for i in range(self.count):
            p = int(self.layout.npad[i])
            descs[i].W = pb + 4 * int(self.layout.offset[i])
            descs[i].snap = sb + int(self.layout.snap_offset[i])
            descs[i].st = stb + i * _lib.STATUS_BYTES
            descs[i].npad = p
            descs[i].tile = self.tiles.get(p, 128)
            descs[i].tri = int(self.tri[i])
            descs[i].snap_bytes = int(self.layout.snap_size[i])
            descs[i].persistent = int(self.persistent_mask[i])
        h = ctypes.c_void_p()
        _lib.call("rbg_batch_create", descs, self.count, ctypes.byref(h))
        self._h = h
        self.descs = descs
        self.captured = False
        self.stream.wait_stream(t.cuda.current_stream())  # the pool's zero fill

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.load().rbg_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        buf = (ctypes.c_int64 * 10)()
        _lib.call("rbg_batch_info", self._h, buf)
        keys = ["count", "steps", "launches", "panel_items", "tiles128", "tiles64", "chunks", "captured",
                "persistent_components", "persistent_items"]
        return dict(zip(keys, [int(x) for x in buf]))

    # -- inputs -------------------------------------------------------------
    def matrix(self, i: int):
        """Padded device view of component i (npad x npad)."""
        p = int(self.layout.npad[i])
        o = int(self.layout.offset[i])
        return self.pool[o:o + p * p].view(p, p)

    def result(self, i: int):
        n = int(self.layout.sizes[i])
        return self.matrix(i)[:n, :n]

    def load_dense(self, mats: Sequence) -> None:
        """Copy component matrices (numpy or device tensors, n x n int32) into the pool."""
        t = _lib.torch()
        with t.cuda.stream(self.stream):
            self.pool.zero_()
            for i, A in enumerate(mats):
                n = int(self.layout.sizes[i])
                src = A if isinstance(A, t.Tensor) else t.from_numpy(np.ascontiguousarray(A, dtype=np.int32))
                self.matrix(i)[:n, :n].copy_(src.to(device=self.pool.device, dtype=t.int32), non_blocking=False)

    def load_edges(self, diag: Sequence[np.ndarray], edges: Sequence[tuple]) -> None:
        """Build every component in the pool from its diagonal and (rows, cols,
        values) edge lists with one device scatter."""
        t = _lib.torch()
        idx, val = [], []
        for i in range(self.count):
            o, p = int(self.layout.offset[i]), int(self.layout.npad[i])
            d = np.asarray(diag[i], dtype=np.int64)
            r = np.arange(d.size, dtype=np.int64)
            idx.append(o + r * p + r)
            val.append(d)
            er, ec, ev = edges[i]
            idx.append(o + np.asarray(er, dtype=np.int64) * p + np.asarray(ec, dtype=np.int64))
            val.append(np.asarray(ev, dtype=np.int64))
        with t.cuda.stream(self.stream):
            ii = t.from_numpy(np.concatenate(idx)).to(self.pool.device)
            vv = t.from_numpy(np.concatenate(val).astype(np.int32)).to(self.pool.device)
            self.pool.zero_()
            self.pool.index_put_((ii,), vv)
        self.stream.synchronize()

    # -- execution ----------------------------------------------------------
    def run(self, graph: bool = True) -> None:
        """Close every component in place.  Asynchronous on self.stream: read
        the pool from another stream only after self.stream.synchronize() (or
        statuses() / outcomes(), which synchronise)."""
        sh = _lib.stream_handle(self.stream)
        if graph:
            if not self.captured:
                _lib.call("rbg_batch_capture", self._h, sh)
                self.captured = True
            _lib.call("rbg_batch_replay", self._h, sh)
        else:
            _lib.call("rbg_batch_run", self._h, sh)

    def statuses(self) -> List[_lib.RbgStatus]:
        t = _lib.torch()
        with t.cuda.stream(self.stream):
            self.status_host.copy_(self.status, non_blocking=True)
        self.stream.synchronize()
        raw = self.status_host.numpy().tobytes()
        S = _lib.STATUS_BYTES
        return [_lib.RbgStatus.from_buffer_copy(raw[i * S:(i + 1) * S]) for i in range(self.count)]

    def outcomes(self, source: Optional[Callable[[int], object]] = None) -> List[FwOutcome]:
        """Per-component outcome (FwOutcome, as the single-matrix path returns).

        Components flagged for the general kernels are closed again from their
        input, ``source(i)`` (an n x n device tensor), and copied back into the
        pool; without ``source`` such a component raises.
        """
        outs = []
        et = elem_type(np.int32)
        for i, s in enumerate(self.statuses()):
            p = int(self.layout.npad[i])
            if s.flags & (_lib.F_NEED_GENERAL | _lib.F_PRECOND_FAIL) or s.cycle_count:
                if source is None:
                    raise RuntimeError(f"component {i} needs the general kernels and no input source was given")
                W, out = fw_closure(source(i), et)
                n = int(self.layout.sizes[i])
                self.matrix(i)[:n, :n].copy_(W)
                outs.append(out)
            else:
                outs.append(FwOutcome(s, p, "fast"))
        return outs


def transitive_closure_batch(mats: Sequence, tiles: Optional[Dict[int, int]] = None, tri=0):
    """Close a list of int32 component matrices together.

    Returns a list of TransitiveClosure (numpy W, diameter) in input order,
    raising the single-matrix closure's errors (core style) for the first
    component that has one.  ``tri`` as for BatchedClosure (pedigrees
    numbered parents first are lower triangular: tri=1).
    """
    from .core import raise_fw_errors
    from .types import TransitiveClosure
    t = _lib.torch()
    mats = [np.asarray(A) for A in mats]
    for A in mats:
        if A.ndim != 2 or A.shape[0] != A.shape[1]:
            raise ValueError("components must be square matrices")
    b = BatchedClosure([A.shape[0] for A in mats], tiles, tri=tri)
    try:
        b.load_dense(mats)
        b.run(graph=False)
        outs = b.outcomes(lambda i: _lib.to_device(mats[i].astype(np.int32)))
        res = []
        et = elem_type(np.int32)
        for i, out in enumerate(outs):
            raise_fw_errors(out, et, "core")
            res.append(TransitiveClosure(b.result(i).cpu().numpy(), msb(out.status.diameter_max)))
        t.cuda.synchronize()
        return res
    finally:
        b.close()
