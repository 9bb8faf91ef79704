"""Closure serving: a request loop with continuous batching over GPU slots.

Requests are either a new pedigree to close (dense int32 below 4096 vertices,
sparse CSR int64 from 4096 up) or the insertion of parent edges into a graph
the server already holds closed.  The loop

  * admits a waiting request as soon as a slot (a worker thread with its own
    HIP stream) is free and its HBM footprint fits the budget, evicting idle
    resident closures least recently used first when it does not;
  * keeps every closed graph resident on the device under its key, so an
    insertion is an incremental relational-composition update of the stored
    closure (core/redblack.py:74-97 / relational_composition.c.in:71-137:
    dense rank-1 AVOS update, or its sparse form through AVOS SpGEMM);
  * lets the work of all slots overlap on the GPU, and keeps the requests of
    one key in arrival order (an insertion waits for the closure or the
    insertion before it);
  * records per-request arrival, start and finish times for p50/p99 latency
    and throughput.

Each request's result depends only on its key's earlier requests, never on
what else is in flight: the same trace served with one slot, request by
request, gives identical closures.
"""
from __future__ import annotations

import hashlib
import math
import queue
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import _lib

DENSE_LIMIT = 4096  # graphs below this many vertices are served dense (int32)


@dataclass
class Request:
    rid: int
    kind: str                 # "close" | "insert"
    arrival: float            # seconds after the trace starts
    key: str                  # name of the resident graph
    graph: object = None      # close: dense np.ndarray (n x n) or scipy CSR
    edges: Optional[list] = None  # insert: [(alpha, beta, relationship), ...]


@dataclass
class Result:
    rid: int
    kind: str
    key: str
    n: int
    arrival: float
    start: float = 0.0
    finish: float = 0.0
    error: Optional[str] = None
    diameter: Optional[int] = None
    entries: int = 0          # closure entries produced (n^2 dense, nnz sparse)
    digest: Optional[str] = None

    @property
    def latency(self) -> float:
        return self.finish - self.arrival


@dataclass
class Resident:
    key: str
    kind: str                 # "dense" | "sparse"
    n: int
    value: object             # dense: torch n x n; sparse: DevCSR
    nbytes: int
    last_used: float = 0.0
    busy: bool = False


def _dense_bytes(n: int) -> int:
    from .dense import padded_order
    p = padded_order(n)
    # input copy, padded work matrix, snapshot workspace (four snapshot
    # buffers and the persistent closure's counters)
    return 4 * n * n + 4 * p * p + 8 * 4 * 32 * p + 4 * (p // 64) ** 2


def _sparse_close_bytes(nnz: int, n: int) -> int:
    # input + canonical copy, workspace, the initial row arena (48 entries per row of A), output
    return 2 * 12 * nnz + 120 * n + 48 * 12 * (nnz + n) + 12 * 64 * n


def _digest(value, kind: str) -> str:
    h = hashlib.sha256()
    if kind == "dense":
        h.update(value.contiguous().cpu().numpy().tobytes())
    else:
        for a in (value.indptr, value.indices, value.data):
            h.update(a.cpu().numpy().tobytes())
    return h.hexdigest()


class ClosureServer:
    """Serve a trace of requests on the current GPU.

    hbm_budget  bytes of device memory the server may hold (residents plus
                the working memory of requests in flight)
    slots       requests that may be on the GPU at once (one stream each)
    digests     record a SHA-256 of every result (for comparisons)
    """

    def __init__(self, hbm_budget: int = 32 << 30, slots: int = 4, digests: bool = False):
        t = _lib.torch()
        _lib.device()
        self.budget = int(hbm_budget)
        self.slots = int(slots)
        self.digests = digests
        self.streams = [t.cuda.Stream() for _ in range(self.slots)]
        self.residents: Dict[str, Resident] = {}
        self.used = 0
        self.peak = 0
        self.evictions = 0
        self._done: "queue.Queue" = queue.Queue()
        self._jobs = [queue.Queue() for _ in range(self.slots)]
        self._workers = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(self.slots)]
        for w in self._workers:
            w.start()

    # -- footprints -----------------------------------------------------------
    def footprint(self, req: Request) -> int:
        if req.kind == "close":
            g = req.graph
            if isinstance(g, np.ndarray):
                return _dense_bytes(g.shape[0])
            return _sparse_close_bytes(int(g.nnz), g.shape[0])
        r = self.residents.get(req.key)
        return 2 * r.nbytes if r is not None else 0

    # -- worker side ------------------------------------------------------------
    def _worker(self, slot: int):
        t = _lib.torch()
        stream = self.streams[slot]
        while True:
            job = self._jobs[slot].get()
            if job is None:
                return
            req, res, resident = job
            res.start = time.perf_counter() - self._t0
            try:
                with t.cuda.stream(stream):
                    value = self._execute(req, res, resident, stream)
                stream.synchronize()
            except Exception as e:  # reported per request
                value = None
                res.error = f"{type(e).__name__}: {e}"
            res.finish = time.perf_counter() - self._t0
            self._done.put((slot, req, res, value))

    def _execute(self, req: Request, res: Result, resident: Optional[Resident], stream):
        from . import dense
        from . import sparse_dev as sd
        from ._dtypes import elem_type
        from .core import raise_fw_errors
        et = elem_type(np.int32)
        if req.kind == "close":
            g = req.graph
            if isinstance(g, np.ndarray):
                Wd = _lib.torch().from_numpy(np.ascontiguousarray(g, dtype=np.int32)).to(_lib.device())
                W, out = dense.fw_closure(Wd, et, stream=stream)
                raise_fw_errors(out, et, "core")
                W = W.contiguous()
                res.diameter, res.entries = out.diameter, g.shape[0] ** 2
                if self.digests:
                    res.digest = _digest(W, "dense")
                return ("dense", W)
            A = sd.canonicalize(sd.from_scipy(g, np.int64))
            C, info = sd.dag_closure(A, stream=stream)
            if C is None:
                raise ValueError(f"graph contains a cycle (vertex {info.cycle_vertex})")
            res.diameter, res.entries = info.diameter_candidates, info.nnz
            if self.digests:
                res.digest = _digest(C, "sparse")
            return ("sparse", C)
        # insertion into a resident closure
        if resident is None:
            raise KeyError(f"no resident graph {req.key!r}")
        value = resident.value
        if resident.kind == "dense":
            value, st = dense.edge_compose_many(value, req.edges, et, stream=stream)
            if st.flags & _lib.F_CYCLE:
                raise ValueError("Relational composition would result in a cycle.")
            if st.events:
                raise OverflowError("Avos product results in an overflow during edge composition")
        else:
            value = sd.edge_insert_many(value, req.edges, stream=stream)
        res.entries = resident.n ** 2 if resident.kind == "dense" else value.nnz
        if self.digests:
            res.digest = _digest(value, resident.kind)
        return (resident.kind, value)

    # -- loop -------------------------------------------------------------------
    def _evict_for(self, need: int, keep: set) -> bool:
        while self.used + need > self.budget:
            idle = [r for r in self.residents.values() if not r.busy and r.key not in keep]
            if not idle:
                return False
            victim = min(idle, key=lambda r: r.last_used)
            del self.residents[victim.key]
            self.used -= victim.nbytes
            self.evictions += 1
        return True

    def serve(self, requests: Sequence[Request], realtime: bool = True) -> List[Result]:
        """Run the trace.  With realtime, a request becomes visible at its
        arrival time; otherwise every request is visible at once (requests
        still start in arrival order per key)."""
        t = _lib.torch()
        reqs = sorted(requests, key=lambda r: (r.arrival, r.rid))
        future = deque(reqs)
        waiting: List[Request] = []
        results: Dict[int, Result] = {}
        free = list(range(self.slots))
        inflight: Dict[int, tuple] = {}       # slot -> (req, reserved bytes)
        busy_keys: set = set()
        self._t0 = time.perf_counter()
        remaining: Dict[str, int] = {}        # key -> requests of that key not finished yet
        for r in reqs:
            remaining[r.key] = remaining.get(r.key, 0) + 1
        self.over_budget = 0
        while future or waiting or inflight:
            now = time.perf_counter() - self._t0
            while future and (not realtime or future[0].arrival <= now):
                waiting.append(future.popleft())
            # completions
            while True:
                try:
                    slot, req, res, value = self._done.get_nowait()
                except queue.Empty:
                    break
                _, reserved = inflight.pop(slot)
                free.append(slot)
                busy_keys.discard(req.key)
                self.used -= reserved
                remaining[req.key] -= 1
                old = self.residents.get(req.key)
                if value is not None:
                    kind, v = value
                    nbytes = (4 * int(v.numel()) if kind == "dense" else
                              12 * v.nnz + 8 * (v.n_rows + 1))
                    if old is not None:
                        self.used -= old.nbytes
                    self.residents[req.key] = Resident(req.key, kind, res.n, v, nbytes, res.finish)
                    self.used += nbytes
                elif old is not None:
                    old.busy = False
                    old.last_used = res.finish
                if req.key in self.residents:
                    self.residents[req.key].busy = False
                results[req.rid] = res
            # admissions: FIFO, skipping requests whose key is busy or whose
            # footprint does not fit yet (later keys may go first)
            blocked_keys = set()
            for req in list(waiting):
                if not free:
                    break
                if req.key in busy_keys or req.key in blocked_keys:
                    blocked_keys.add(req.key)
                    continue
                if req.kind == "insert" and req.key not in self.residents:
                    if any(r.key == req.key and r.kind == "close" for r in waiting if r is not req) or \
                            any(q.key == req.key for q, _ in inflight.values()):
                        blocked_keys.add(req.key)
                        continue
                need = self.footprint(req)
                # residents with requests still to come stay, unless nothing
                # is in flight to free memory (then the oldest idle go, and a
                # request that still does not fit runs over budget)
                keep = busy_keys | {req.key} | {k for k, c in remaining.items() if c > 0}
                if not self._evict_for(need, keep):
                    if inflight:
                        blocked_keys.add(req.key)
                        continue
                    if not self._evict_for(need, busy_keys | {req.key}):
                        self.over_budget += 1
                slot = free.pop()
                waiting.remove(req)
                busy_keys.add(req.key)
                resident = self.residents.get(req.key)
                if resident is not None:
                    resident.busy = True
                n = (req.graph.shape[0] if req.kind == "close" else (resident.n if resident else 0))
                res = Result(req.rid, req.kind, req.key, int(n), req.arrival)
                self.used += need
                self.peak = max(self.peak, self.used)
                inflight[slot] = (req, need)
                self._jobs[slot].put((req, res, resident))
            if inflight or waiting or future:
                # wait for a completion or the next arrival
                timeout = 0.002
                if future and realtime and not inflight:
                    timeout = max(0.0, min(0.05, future[0].arrival - (time.perf_counter() - self._t0)))
                try:
                    item = self._done.get(timeout=timeout)
                    self._done.put(item)
                except queue.Empty:
                    pass
        t.cuda.synchronize()
        return [results[r.rid] for r in reqs]

    def close(self):
        for q in self._jobs:
            q.put(None)
        for w in self._workers:
            w.join(timeout=10)


# ---------------------------------------------------------------------------
# synthetic trace
# ---------------------------------------------------------------------------
def synthetic_trace(count: int = 200, rate: float = 50.0, p_insert: float = 0.35, lo: int = 32, hi: int = 16384,
                    seed: int = 0, max_edges: int = 64) -> List[Request]:
    """Poisson arrivals at ``rate`` per second.  A request closes a new
    pedigree (size log-uniform in [lo, hi], layers = ceil(log2 n), parents
    known with p=0.8 from the previous 1-3 layers; dense int32 below 4096
    vertices, CSR int64 from 4096) or, with probability p_insert once some
    graph exists, inserts 1..max_edges parent edges into an earlier graph:
    child alpha in the later half, parent beta < alpha of the sex the
    relationship needs (2 father / 3 mother), so no edge closes a cycle."""
    from . import synth
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    reqs: List[Request] = []
    graphs: List[tuple] = []  # (key, n, is_red)
    for rid in range(count):
        t += rng.exponential(1.0 / rate)
        if graphs and rng.random() < p_insert:
            key, n, is_red = graphs[int(rng.integers(0, len(graphs)))]
            k = int(rng.integers(1, max_edges + 1))
            edges = []
            for _ in range(k):
                a = int(rng.integers(max(1, n // 2), n))
                rel = 2 if rng.random() < 0.5 else 3
                cand = np.nonzero(is_red[:a] == (rel == 2))[0]
                if cand.size == 0:
                    continue
                b = int(cand[int(rng.integers(0, cand.size))])
                edges.append((a, b, rel))
            if not edges:
                edges = [(n - 1, 0, 2 if is_red[0] else 3)]
            reqs.append(Request(rid, "insert", t, key, edges=edges))
            continue
        n = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        spec = synth.component_spec(n)
        gseed = seed * 1_000_003 + rid
        is_red, child, parent, value = synth.pedigree_edges(spec, gseed)
        if n < DENSE_LIMIT:
            g = synth.dense_pedigree(spec, gseed)
        else:
            g = synth.csr_pedigree(spec, gseed)
        key = f"g{rid}"
        graphs.append((key, n, is_red))
        reqs.append(Request(rid, "close", t, key, graph=g))
    return reqs


def warmup_trace(hi: int = 16384, max_edges: int = 64, seed: int = 0) -> List[Request]:
    """Requests of the largest shapes a trace from synthetic_trace(hi=hi) can
    hold: the largest dense graph and a graph of ``hi`` vertices, closed and
    then given ``max_edges`` insertions each, all arriving at once.  Serving
    them first leaves the device memory pools holding blocks of every size the
    trace will ask for, so no measured request waits for a fresh allocation."""
    from . import synth
    rng = np.random.Generator(np.random.PCG64(seed))
    reqs: List[Request] = []
    for i, n in enumerate((DENSE_LIMIT - 1, hi)):
        spec = synth.component_spec(n)
        gseed = 7_000_003 + seed * 101 + i
        is_red, _, _, _ = synth.pedigree_edges(spec, gseed)
        g = synth.dense_pedigree(spec, gseed) if n < DENSE_LIMIT else synth.csr_pedigree(spec, gseed)
        key = f"warm{i}"
        reqs.append(Request(2 * i, "close", 0.0, key, graph=g))
        edges = []
        for _ in range(max_edges):
            a = int(rng.integers(max(1, n // 2), n))
            rel = 2 if rng.random() < 0.5 else 3
            cand = np.nonzero(is_red[:a] == (rel == 2))[0]
            if cand.size:
                edges.append((a, int(cand[int(rng.integers(0, cand.size))]), rel))
        reqs.append(Request(2 * i + 1, "insert", 0.0, key, edges=edges))
    return reqs


def summarize(results: Sequence[Result], makespan: Optional[float] = None) -> dict:
    ok = [r for r in results if r.error is None]
    lat = np.array([r.latency for r in ok]) * 1e3 if ok else np.zeros(1)
    span = makespan if makespan is not None else (max(r.finish for r in results) if results else 0.0)
    closes = [r for r in ok if r.kind == "close"]
    inserts = [r for r in ok if r.kind == "insert"]

    def pct(xs, q):
        return float(np.percentile(np.array([x.latency for x in xs]) * 1e3, q)) if xs else None
    return {
        "requests": len(results), "errors": len(results) - len(ok), "makespan_s": span,
        "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
        "close_p50_ms": pct(closes, 50), "close_p99_ms": pct(closes, 99),
        "insert_p50_ms": pct(inserts, 50), "insert_p99_ms": pct(inserts, 99),
        "closures_per_s": len(closes) / span if span else 0.0,
        "requests_per_s": len(ok) / span if span else 0.0,
        "entries_per_s": sum(r.entries for r in ok) / span if span else 0.0,
    }
