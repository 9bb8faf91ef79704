"""The persistent lower triangular closure's schedule, checked on the CPU.

rbg_fw_flow_schedule lists the work items in ticket order and
rbg_fw_flow_waits each item's wait conditions, both with the code the kernel
itself runs (flow_item, flow_cond in fw_kernels.hpp).  These tests check
that the list holds every item once; that every condition names only items
of smaller tickets (so no wait cycle can form, whichever workgroups run);
and that no two items the waits leave unordered write the same 32 x 32
block or a block the other reads -- apart from the reads a tile makes right
of its pivot block, which it neither stores nor relaxes (a lower triangular
step's products there are absent) -- or share a snapshot buffer for
different pivot blocks.
"""
import ctypes

import numpy as np
import pytest

from redblackgraph_amd import _lib

BLK = 32
MAXW = 10


def _edge(n, elem, edge=0):
    """The tile edge: `edge`, or with 0 the single-matrix closure's choice."""
    return edge or (128 if elem == 4 and n > 2048 else 64)


def _schedule(n, elem, edge=0):
    L = _lib.load()
    total = L.rbg_fw_flow_schedule(n, elem, edge, None, 0)
    assert total > 0
    buf = np.zeros(4 * total, dtype=np.int32)
    assert L.rbg_fw_flow_schedule(n, elem, edge, buf.ctypes.data_as(ctypes.c_void_p), buf.size) == total
    w = np.zeros(2 * MAXW * total, dtype=np.int64)
    assert L.rbg_fw_flow_waits(n, elem, edge, w.ctypes.data_as(ctypes.c_void_p), w.size) == total
    return [tuple(int(v) for v in row) for row in buf.reshape(total, 4)], w.reshape(total, MAXW, 2)


def _layout(n, elem, edge=0):
    nb, edge = n // BLK, _edge(n, elem, edge)
    nt = n // edge
    lo = [K * BLK // edge for K in range(nb)]
    return nb, nt, edge, lo


def _ncol(nb, K):
    return (nb - K) // 2


def _panel_item(nb, K, idx):
    """Item idx of P(K): ("c", g) for the column panel's rows from block
    K + 1 + 2 g (two blocks each), ("r", J) for row panel (K, J).  Order:
    C(K, 0), R(K, 0 .. K - 1), C(K, 1 ..)."""
    c = 1 if _ncol(nb, K) > 0 else 0
    if c and idx == 0:
        return "c", 0
    if idx < K + c:
        return "r", idx - c
    return "c", idx - K


def _all_items(n, elem, edge=0):
    nb, nt, edge, lo = _layout(n, elem, edge)
    items = set()
    for K in range(nb):
        items |= {(0, K, idx, 0) for idx in range(K + _ncol(nb, K))}
        items |= {(1, K, I, J) for I in range(lo[K], nt) for J in range(lo[K] + 1)}
    return items


def _deps(n, elem, sched, waits, edge=0):
    """Per ticket: the tickets its conditions wait for (a condition on a
    counter holds once every item it counts has finished)."""
    nb, nt, edge, lo = _layout(n, elem, edge)
    ticket = {it: x for x, it in enumerate(sched)}
    panels_of, tiles_of = {}, {}
    for x, it in enumerate(sched):
        (panels_of if it[0] == 0 else tiles_of).setdefault(it[1], []).append(x)
    base_t, base_s, base_f = 4 + nb, 4 + 2 * nb, 4 + 2 * nb + nt * nt
    out = []
    for x in range(len(sched)):
        dx = []
        for word, target in waits[x]:
            if word < 0:
                continue
            if word < base_t:  # finished items of P(K)
                K = word - 4
                assert target == len(panels_of[K])
                dx += panels_of[K]
            elif word < base_s:  # finished tiles of T(K)
                K = word - base_t
                assert target == len(tiles_of[K])
                dx += tiles_of[K]
            elif word < base_f:  # tile (I, J) done through step target - 1
                I, J = divmod(word - base_s, nt)
                dx.append(ticket[(1, target - 1, I, J)])
            else:  # panel item
                K, idx = divmod(word - base_f, nb)
                assert target == 1
                dx.append(ticket[(0, K, idx, 0)])
        out.append(dx)
    return out


# (n, elem, edge): the single-matrix closure's shapes (edge 0) and the
# batched closure's 128-tiles at small orders
@pytest.mark.parametrize("n,elem,edge", [(128, 8, 0), (512, 4, 0), (1152, 8, 0), (2176, 4, 0), (4096, 8, 0),
                                         (8192, 4, 0), (128, 4, 128), (384, 4, 128), (1024, 4, 128),
                                         (640, 4, 64)])
def test_flow_tickets_cover_items_and_wait_backwards(n, elem, edge):
    sched, waits = _schedule(n, elem, edge)
    items = _all_items(n, elem, edge)
    assert len(sched) == len(items) and set(sched) == items
    # round of an item: K for P(K), K + 1 for T(K); every wait names an
    # earlier round, so no two items of a round ever wait for each other
    rnd = [it[1] + (it[0] == 1) for it in sched]
    for x, dx in enumerate(_deps(n, elem, sched, waits, edge)):
        assert all(d < x for d in dx), (sched[x], [sched[d] for d in dx if d >= x])
        assert all(rnd[d] < rnd[x] for d in dx), (sched[x], [sched[d] for d in dx if rnd[d] >= rnd[x]])
    # within each step's tiles, block K + 2's tile row comes first
    nb, nt, edge, lo = _layout(n, elem, edge)
    for K in range(nb - 2):
        tiles = [it for it in sched if it[0] == 1 and it[1] == K]
        assert all(t[2] == lo[K + 2] for t in tiles[: lo[K] + 1])


def _blocks(item, nb, edge):
    """(writes, reads, loose reads) of an item as sets of 32 x 32 blocks
    (row block, column block); loose reads are read but neither relaxed nor
    stored.  (The diagonal blocks are closed before the launch.)"""
    role, K, a, b = item
    if role == 0:
        kind, v = _panel_item(nb, K, a)
        if kind == "r":
            blk = {(K, v)}
        else:
            s0 = K + 1 + 2 * v
            blk = {(s, K) for s in range(s0, min(s0 + 2, nb))}
        return blk, set(blk), set()
    per = edge // BLK
    rows = [r for r in range(a * per, a * per + per) if r not in (K, K + 1)]
    cols = range(b * per, b * per + per)
    w = {(r, c) for r in rows for c in cols if c <= K}
    loose = {(r, c) for r in rows for c in cols if c > K}
    return w, set(w), loose


def _snap_blocks(item, nb, edge):
    """Snapshot segments an item writes and reads: (step, "c" | "r", block)
    for step K's column snapshots of a block's rows or row snapshots of its
    columns.  The diagonal blocks' come from their own buffer, written before
    the launch; the others are written by P(K) and read by T(K) and by row
    panels of P(K + 1)."""
    role, K, a, b = item
    if role == 0:
        kind, v = _panel_item(nb, K, a)
        if kind == "r":  # row panel (K, v): its strip reads step K - 1's
            r = {(K - 1, "c", K)} | ({(K - 1, "r", v)} if v < K - 1 else set())
            return {(K, "r", v)}, r
        s0 = K + 1 + 2 * v
        return {(K, "c", s) for s in range(s0, min(s0 + 2, nb))}, set()
    per = edge // BLK
    r = {(K, "c", s) for s in range(a * per, a * per + per) if s > K}
    r |= {(K, "r", s) for s in range(b * per, b * per + per) if s < K}
    return set(), r


@pytest.mark.parametrize("n,elem,edge", [(256, 8, 0), (1024, 4, 0), (2048, 8, 0), (2560, 4, 0), (128, 4, 128),
                                         (256, 4, 128), (768, 4, 128), (1152, 4, 128)])
def test_flow_unordered_items_touch_disjoint_blocks(n, elem, edge):
    sched, waits = _schedule(n, elem, edge)
    nb, nt, edge, lo = _layout(n, elem, edge)
    deps = _deps(n, elem, sched, waits, edge)
    anc = [0] * len(sched)  # ancestors as bit sets over tickets
    for x, dx in enumerate(deps):
        m = 0
        for d in dx:
            m |= anc[d] | (1 << d)
        anc[x] = m

    def ordered(x, y):
        return (anc[max(x, y)] >> min(x, y)) & 1

    def check(writers, touchers):
        for key, ws in writers.items():
            for x in ws:
                for y in touchers.get(key, ()):
                    assert y == x or ordered(x, y), (key, sched[x], sched[y])

    writers, touchers = {}, {}
    for x, it in enumerate(sched):
        w, r, _ = _blocks(it, nb, edge)
        for blk in w:
            writers.setdefault(blk, []).append(x)
        for blk in w | r:
            touchers.setdefault(blk, []).append(x)
    check(writers, touchers)
    # snapshot buffers: step K's live in buffer K mod 4; a write of one step's
    # segment must be ordered with every read or write of that buffer's
    # segment for any step (the same step: its readers; others: the ring)
    writers, touchers = {}, {}
    for x, it in enumerate(sched):
        w, r = _snap_blocks(it, nb, edge)
        for (K, kind, s) in w:
            writers.setdefault((K % 4, kind, s), []).append(x)
        for (K, kind, s) in w | r:
            touchers.setdefault((K % 4, kind, s), []).append(x)
    check(writers, touchers)
