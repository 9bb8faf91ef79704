#!/usr/bin/env python3
"""AVOS closure throughput on one MI355X.

Default workload: dense Floyd-Warshall closure of the seeded N=8192 int32
pedigree (24 layers, parents known with p=0.75 from the previous 1-4 layers).
A step is one full closure of that matrix, input already resident in HBM,
through the same entry point users call (redblackgraph_amd.dense.fw_closure:
copy, fast-path precheck, closure, status read-back).

Prints one JSON line (the bench contract) with:
  value          closure entries/s = N^2 / seconds per closure
  roofline       the dominant kernel (for the lower triangular N=8192 input
                 the persistent closure k_fw_flow<int>: every pivot step's
                 panels and phase-3 tiles in one launch, work items waiting
                 on each other): integer VALU ops it must execute (2.5 per
                 relaxation) over its average launch duration, measured with
                 hipEvents on its launch stream
                 (recorded on every --prof-every-th timed step, 4 by default,
                 so that the events' own stalls stay out of most steps)
  cpu_baseline   the oracle's single-core naive loop (a C restatement of the
                 reference's warshall.c.in) on a bounded sample, scaled to the
                 workload by the N^3 law
The default (--workload all) nests the rest of the metric under
"sub_results": the sparse 1M-vertex closure, the batched 4096-component
closure against both per-graph paths, the N=4096 int64 closure, the
served stream, the dense N=8192 product and the canonical ordering of the
1M closure.  Single workloads: --workload dense8192, dense4096_i64,
dense512, sparse1m, batched, serve, order1m, matmul8192, hbmcopy.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AVOS closure entries/sec, dense N=8192 & sparse 1M-vertex; batched vs plain"
PEAK_VALU_OPS = 256 * 64 * 2.4e9  # int32 VALU lane-ops/s (256 CU x 64 lanes x 2.4 GHz)
PEAK_HBM = 8.0e12
OPS_PER_RELAX = {"int32": 2.5, "int64": 9.0}
MM_OPS_PER_TERM = 1.5  # dense fast product (k_mm_fast<int, 8>)
FW_BLK = 32


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def _read_traffic(name: str):
    """HBM bytes per launch from a committed rocprofv3 PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and not line.startswith("#"):
                k, v = line.split(":", 1)
                try:
                    vals[k.strip()] = float(v.split()[0])
                except (ValueError, IndexError):
                    pass
    if "traffic_bytes_per_launch" in vals:
        return vals["traffic_bytes_per_launch"]
    return None


class _one_core:
    """Pins the calling thread to one core of its affinity set for the CPU
    baseline (the reference has no CPU parallelism), restoring the set after."""

    def __enter__(self):
        self.saved = os.sched_getaffinity(0)
        self.core = min(self.saved)
        os.sched_setaffinity(0, {self.core})
        return self

    def __exit__(self, *exc):
        os.sched_setaffinity(0, self.saved)
        return False

    def host(self) -> str:
        return (f"host CPU: {_cpu_model()}; nproc {os.cpu_count()}, affinity set of {len(self.saved)} cores, "
                f"run pinned to core {self.core}")


def cpu_baseline_dense(n_sample: int, n_target: int, dtype):
    """Single-core naive FW (oracle restatement of warshall.c.in) on an
    n_sample pedigree; relaxations/s scaled to entries/s at n_target."""
    import numpy as np
    from oracle import oracle
    from redblackgraph_amd import synth
    A = synth.dense_pedigree(synth.PedigreeSpec(n_sample, max(4, round(24 * n_sample / 8192)), 0.75, 4), 1,
                             dtype=dtype)
    oracle.lib()
    with _one_core() as pin:
        t0 = time.perf_counter()
        oracle.warshall(A)
        dt = time.perf_counter() - t0
    rate = n_sample ** 3 / dt  # relaxations per second
    return {
        "value": rate / n_target,  # entries/s = N^2 / (N^3 / rate)
        "unit": "entries/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"oracle naive k-i-j FW (C restatement of warshall.c.in, gcc -O2, 1 thread) on a "
                   f"seeded N={n_sample} {np.dtype(dtype).name} pedigree: {dt:.2f} s = {rate:.3e} relaxations/s; "
                   f"value = that rate / {n_target} (N^3 scaling to N={n_target}: the naive loop runs all N^3 "
                   f"relaxations whatever the input); {pin.host()}"),
    }


def _sampled_steps(args) -> int:
    """Timed steps whose kernels were bracketed by timing events (every
    prof_every-th step, starting with the first)."""
    return (args.steps + args.prof_every - 1) // args.prof_every


def _prof_read(kind: int):
    """(summed ms, launches, recorded work) of one timed launch kind since the
    last read (rbg_prof_read_work)."""
    import ctypes
    from redblackgraph_amd import _lib
    ms, nl, wk = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double()
    _lib.call("rbg_prof_read_work", kind, ctypes.byref(ms), ctypes.byref(nl), ctypes.byref(wk))
    return ms.value, nl.value, wk.value


def bench_dense(args, n: int, dtype_name: str):
    import numpy as np
    from redblackgraph_amd import _lib, dense, synth
    from redblackgraph_amd._dtypes import elem_type
    t = _lib.torch()
    dtype = np.int32 if dtype_name == "int32" else np.int64
    spec = synth.DENSE_8192 if n == 8192 else synth.PedigreeSpec(n, 24, 0.75, 4)
    A = synth.dense_pedigree(spec, args.seed, dtype=dtype)
    et = elem_type(dtype)
    Ad = _lib.to_device(A)
    L = _lib.load()
    sched = getattr(args, "dense_schedule", "auto")
    for _ in range(args.warmup):
        W, out = dense.fw_closure(Ad, et, schedule=sched)
    t.cuda.synchronize()
    L.rbg_fw_prof_enable(1)
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        L.rbg_prof_enable(2 if i % args.prof_every == 0 else 0)
        W, out = dense.fw_closure(Ad, et, schedule=sched)
    t.cuda.synchronize()
    t1 = time.perf_counter()
    sampled = _sampled_steps(args)
    # panel, 128-tiles, 64-tiles, persistent closure: (ms, launches, relaxations)
    prof = {k: _prof_read(k) for k in (0, 1, 5, 6)}
    L.rbg_fw_prof_enable(0)
    if out.mode != "fast" or out.status.cycle_count or out.status.events:
        raise RuntimeError(f"unexpected closure outcome: mode={out.mode} status={out.status.flags}")
    ms = (t1 - t0) * 1e3 / args.steps
    # host-boundary rate: numpy matrix in, numpy closure out (PCIe both ways)
    t.cuda.synchronize()
    h0 = time.perf_counter()
    Wh = dense.fw_closure(_lib.to_device(A), et, schedule=sched)[0].cpu().numpy()
    pcie_ms = (time.perf_counter() - h0) * 1e3
    del Wh
    npad = out.npad
    # the dominant kernel: for lower triangular inputs the persistent closure
    # (every pivot step's panels and tiles in one launch), else the 128 x 128
    # (int32) / 64 x 64 (int64) phase-3 tile launches; achieved = relaxations
    # its launches executed (the triangular schedule's skipped tiles are not
    # counted, the panels' are) x ops per relaxation / time
    tm, tl, tw = prof[6] if prof[6][1] else prof[1]
    tile_avg_s = tm / max(tl, 1) / 1e3
    achieved = (tw / max(tl, 1)) * OPS_PER_RELAX[dtype_name] / tile_avg_s
    if prof[6][1]:
        kname = "k_fw_flow<int, 8, false, 3>" if dtype_name == "int32" else "k_fw_flow<long, 4, true, 3>"
        traffic = _read_traffic(f"fw_flow_dense{n}_{dtype_name}_pmc.txt")
    elif out.schedule in ("lower", "lower-steps"):
        # the fused look-ahead step: pivot block K's tiles and block K+1's panel in one launch
        kname = "k_fw_step_fast<int, 8, false, 3>" if dtype_name == "int32" else "k_fw_step_fast<long, 4, true, 3>"
        traffic = _read_traffic(f"fw_step_fast_dense{n}_{dtype_name}_pmc.txt")
    else:
        kname = f"k_fw_tile_fast<{'int' if dtype_name == 'int32' else 'long'}>"
        traffic = _read_traffic(f"fw_tile_fast_dense{n}_{dtype_name}_pmc.txt")
    sm, sl, sw = prof[5]
    pm, pl, _ = prof[0]
    roof = {
        "bound": "valu", "kernel": kname,
        "achieved": achieved / 1e12, "peak": PEAK_VALU_OPS / 1e12, "unit": "TOP/s",
        "frac": achieved / PEAK_VALU_OPS,
        "traffic": traffic,
        "ops_per_relaxation": OPS_PER_RELAX[dtype_name],
        "relaxations_per_launch": tw / max(tl, 1),
        "avg_launch_ms": tile_avg_s * 1e3, "launches": tl,
        "schedule": out.schedule,
        "relaxations_executed_per_closure": (tw + sw) / sampled,
        "relaxations_full_schedule": (npad - FW_BLK) * npad * npad,
        "small_tile_launches": sl, "small_tile_avg_ms": sm / max(sl, 1),
        "small_tile_frac": ((sw / sm * 1e3 * OPS_PER_RELAX[dtype_name] / PEAK_VALU_OPS) if sm else None),
        "panel_avg_ms": pm / max(pl, 1),
        "panel_ms_per_closure": pm / sampled,
        "kernel_share_of_step": (tm + sm + pm) / (ms * sampled),
        "timed_steps_sampled": sampled,
    }
    res = {
        "metric": METRIC, "value": n * n / (ms / 1e3), "unit": "entries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": dtype_name, "data": "synthetic",
        "config": {"workload": f"dense Floyd-Warshall closure, N={n} {dtype_name}, seeded layered pedigree "
                               f"({spec.layers} layers, p_parent={spec.p_parent}, lookback {spec.lookback}), seed {args.seed}",
                   "n": n, "diameter": out.diameter, "relaxations": npad ** 3,
                   "host_boundary_ms": pcie_ms, "host_boundary_entries_per_s": n * n / (pcie_ms / 1e3)},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_dense(args.cpu_sample, n, dtype)
    return res


def cpu_baseline_sparse(A, nnz_c: int):
    """Single-core oracle DAG closure (C restatement of _transitive_closure_dag.pyx)
    on the full workload."""
    from oracle import oracle
    oracle.lib()
    with _one_core() as pin:
        t0 = time.perf_counter()
        r = oracle.dag_closure(A, 64)
        dt = time.perf_counter() - t0
    assert r.indptr[-1] == nnz_c
    return {
        "value": nnz_c / dt, "unit": "entries/s", "cores": 1, "kind": "port",
        "sample": (f"oracle sparse DAG closure (C restatement of _transitive_closure_dag.pyx:168-351, gcc -O2, "
                   f"1 thread) on the full 1M-vertex workload: {dt:.2f} s for {nnz_c} closure entries; "
                   f"{pin.host()}"),
    }


def bench_sparse(args):
    import ctypes
    import numpy as np
    from redblackgraph_amd import _lib, synth
    from redblackgraph_amd import sparse_dev as sd
    t = _lib.torch()
    A = synth.csr_pedigree(synth.SPARSE_1M, args.seed)
    Ad = sd.from_scipy(A, np.int64)
    plan = None
    L = _lib.load()
    for _ in range(max(1, args.warmup)):
        C, info = sd.dag_closure(sd.canonicalize(Ad), plan=plan)
    plan = sd.DagPlan(A.shape[0], t.int64, info.arena + 4096, A.nnz)
    C, info = sd.dag_closure(sd.canonicalize(Ad), plan=plan)
    t.cuda.synchronize()
    L.rbg_prof_enable(1)
    t0 = time.perf_counter()
    for i in range(args.steps):
        L.rbg_prof_enable(2 if i % args.prof_every == 0 else 0)
        C, info = sd.dag_closure(sd.canonicalize(Ad), plan=plan)
    t.cuda.synchronize()
    t1 = time.perf_counter()
    sampled = _sampled_steps(args)
    ms_, nl = ctypes.c_double(), ctypes.c_int64()
    _lib.call("rbg_prof_read", 2, ctypes.byref(ms_), ctypes.byref(nl))
    L.rbg_prof_enable(0)
    ms = (t1 - t0) * 1e3 / args.steps
    # first closure of a graph the plan has not seen: no level count, level
    # kinds or output size known in advance (the device memory pool is warm)
    first = []
    for _ in range(3):
        t.cuda.synchronize()
        f0 = time.perf_counter()
        Cf, _ = sd.dag_closure(sd.canonicalize(Ad))
        t.cuda.synchronize()
        first.append((time.perf_counter() - f0) * 1e3)
        del Cf
    first_ms = sorted(first)[1]
    # host-boundary rate: scipy CSR in, scipy CSR out (PCIe both ways)
    t.cuda.synchronize()
    h0 = time.perf_counter()
    Ch, _ = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)), plan=plan)
    Ch = sd.to_scipy(Ch)
    pcie_ms = (time.perf_counter() - h0) * 1e3
    del Ch
    n = A.shape[0]
    clen = np.diff(C.indptr.cpu().numpy())
    rows = np.repeat(np.arange(n), np.diff(A.indptr))
    off = A.indices != rows
    reads = int(clen[A.indices[off]].sum())
    bytes_alg = 12 * A.nnz + 12 * reads + 12 * info.nnz + 8 * (n + 1)
    kernel_s = ms_.value / 1e3 / sampled
    achieved = bytes_alg / kernel_s
    roof = {
        "bound": "hbm", "kernel": "k_lvl_units<long> + k_lvl_pairs<long> + k_lvl_long<long> (level merge)",
        "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": achieved / PEAK_HBM,
        "traffic": _read_traffic("lvl_merge_sparse1m_pmc.txt"),
        "algorithmic_bytes_per_closure": bytes_alg,
        "launches_per_closure": nl.value / sampled, "avg_launch_ms": ms_.value / max(nl.value, 1),
        "kernel_ms_per_closure": kernel_s * 1e3,
        "kernel_share_of_step": kernel_s * 1e3 / ms,
        "timed_steps_sampled": sampled,
    }
    res = {
        "metric": METRIC, "value": info.nnz / (ms / 1e3), "unit": "entries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic",
        "config": {"workload": f"sparse DAG closure, 1M-vertex pedigree (15 layers of ~66,667, p_parent=0.7, "
                               f"parents from the previous layer), int64 values / int32 indices, seed {args.seed}",
                   "n": n, "nnz_in": int(A.nnz), "nnz_closure": info.nnz, "levels": info.nlevels,
                   "diameter": info.diameter_candidates,
                   "primed_plan": "the timed closures reuse the plan of an earlier closure of this graph "
                                  "(level count, level kinds, output size)",
                   "first_closure_ms": first_ms,
                   "first_closure_entries_per_s": info.nnz / (first_ms / 1e3),
                   "host_boundary_ms": pcie_ms,
                   "host_boundary_entries_per_s": info.nnz / (pcie_ms / 1e3)},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_sparse(A, info.nnz)
    return res


def cpu_baseline_batched(sizes, diag, edges, budget_s: float = 15.0):
    """Single-core naive FW (oracle) on the batch's components in order of
    size, smallest first, until about budget_s of CPU time; the measured
    relaxations/s give the whole batch's CPU time by the n^3 law."""
    import numpy as np
    from oracle import oracle
    oracle.lib()
    order = np.argsort(sizes, kind="stable")
    done_relax, spent, used = 0, 0.0, 0
    with _one_core() as pin:
        for i in order[:: max(1, len(order) // 256)]:
            n = int(sizes[i])
            A = np.zeros((n, n), dtype=np.int32)
            r, c, v = edges[i]
            A[r, c] = v
            np.fill_diagonal(A, diag[i])
            t0 = time.perf_counter()
            oracle.warshall(A)
            spent += time.perf_counter() - t0
            done_relax += n ** 3
            used += 1
            if spent > budget_s:
                break
    rate = done_relax / spent
    entries = float(np.sum(np.asarray(sizes, dtype=np.float64) ** 2))
    relax = float(np.sum(np.asarray(sizes, dtype=np.float64) ** 3))
    return {
        "value": entries / (relax / rate), "unit": "entries/s", "cores": 1, "kind": "port",
        "sample": (f"oracle naive FW (C restatement of warshall.c.in, gcc -O2, 1 thread) on {used} of the batch's "
                   f"components (every {max(1, len(order) // 256)}th by size, smallest first): {spent:.2f} s, "
                   f"{rate:.3e} relaxations/s; value = sum n^2 / (sum n^3 / rate); {pin.host()}"),
    }


def _batched_plain_legs(b, sizes, pristine, eager_result):
    """The batch's components closed one graph at a time: by the same kernels
    (a one-component schedule each, on one stream), and by the single-matrix
    entry point.  Returns (plain_ms, identical, single_ms, identical)."""
    import ctypes
    import numpy as np
    from redblackgraph_amd import _lib, dense
    from redblackgraph_amd._dtypes import elem_type
    t = _lib.torch()
    L = _lib.load()
    # plain path: the same kernels launched one graph at a time on one stream
    # (a one-component schedule per graph, built before the clock starts)
    handles = []
    for i in range(b.count):
        one = (_lib.RbgBatchComp * 1)(b.descs[i])
        h = ctypes.c_void_p()
        _lib.call("rbg_batch_create", one, 1, ctypes.byref(h))
        handles.append(h)
    sh = _lib.stream_handle(b.stream)
    with t.cuda.stream(b.stream):
        b.pool.copy_(pristine)
    b.stream.synchronize()
    tp0 = time.perf_counter()
    for h in handles:
        _lib.call("rbg_batch_run", h, sh)
    b.stream.synchronize()
    plain_ms = (time.perf_counter() - tp0) * 1e3
    identical_plain = bool(t.equal(b.pool, eager_result))
    for h in handles:
        L.rbg_batch_destroy(h)
    # the single-matrix entry point, one closure per component (its own panel
    # kernel, a status read per closure)
    et = elem_type(np.int32)
    views = []
    for i, n in enumerate(sizes):
        p, o = int(b.layout.npad[i]), int(b.layout.offset[i])
        views.append(pristine[o:o + p * p].view(p, p)[:int(n), :int(n)])
    t.cuda.synchronize()
    tp0 = time.perf_counter()
    single = [dense.fw_closure(V, et)[0] for V in views]
    t.cuda.synchronize()
    single_ms = (time.perf_counter() - tp0) * 1e3
    identical_single = all(bool(t.equal(single[i], b.result(i))) for i in range(len(sizes)))
    del single
    return plain_ms, identical_plain, single_ms, identical_single


def bench_batched(args):
    import ctypes
    import numpy as np
    from redblackgraph_amd import _lib, dense, synth
    from redblackgraph_amd._dtypes import elem_type
    from redblackgraph_amd.batched import BatchedClosure
    t = _lib.torch()
    sizes, diag, edges = synth.batched_workload(count=args.batch_count, seed=args.seed)
    t_tune0 = time.perf_counter()
    # pedigrees numbered parents first: lower triangular; the schedule is the
    # autotuner's choice for this mix of orders
    b = BatchedClosure(sizes, tri=1, persistent="auto")
    tune_s = time.perf_counter() - t_tune0
    b.load_edges(diag, edges)
    pristine = b.pool.clone()
    t.cuda.synchronize()
    info = b.info()
    L = _lib.load()
    # eager run with per-step tile timing (prof events do not go into graphs)
    L.rbg_prof_enable(1)
    b.run(graph=False)
    b.stream.synchronize()
    tile_ms, tile_launches, relax_exec = _prof_read(4)
    L.rbg_prof_enable(0)
    outs = b.outcomes()
    if any(o.mode != "fast" or o.status.cycle_count or o.status.events for o in outs):
        raise RuntimeError("unexpected component outcome in the batched workload")
    eager_result = b.pool.clone()
    t.cuda.synchronize()  # the clone (current stream) must finish before the batch stream rewrites the pool
    with t.cuda.stream(b.stream):
        for _ in range(args.warmup):
            b.pool.copy_(pristine)
            b.run(graph=True)
    b.stream.synchronize()
    t.cuda.synchronize()
    t0 = time.perf_counter()
    with t.cuda.stream(b.stream):
        for _ in range(args.steps):
            b.pool.copy_(pristine)
            b.run(graph=True)
    b.stream.synchronize()
    t.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    identical_graph = bool(t.equal(b.pool, eager_result))
    if args.no_plain:
        plain_ms = single_ms = None
        identical_plain = identical_single = None
    else:
        plain_ms, identical_plain, single_ms, identical_single = _batched_plain_legs(b, sizes, pristine, eager_result)
    tile_s = tile_ms / 1e3
    achieved = relax_exec * OPS_PER_RELAX["int32"] / tile_s
    entries = float(np.sum(sizes.astype(np.float64) ** 2))
    persistent = info["persistent_components"] == len(sizes)
    roof = {
        "bound": "valu",
        "kernel": ("k_bfw_flow (every component's panel items and tiles in one persistent launch)" if persistent
                   else "k_bfw_tile<8>/<4> (batched phase 3)"),
        "achieved": achieved / 1e12, "peak": PEAK_VALU_OPS / 1e12, "unit": "TOP/s",
        "frac": achieved / PEAK_VALU_OPS,
        "traffic": _read_traffic("bfw_flow_batched4096_pmc.txt" if persistent else "bfw_tile_batched4096_pmc.txt"),
        "ops_per_relaxation": OPS_PER_RELAX["int32"], "relaxations_executed": relax_exec,
        "schedule": ("lower triangular, persistent: every component's work items in one grid" if persistent
                     else "lower triangular (the tiles each pivot block can change)"),
        "tile_ms_per_batch": tile_ms, "tile_launches": tile_launches,
        "kernel_share_of_step": tile_ms / ms,
    }
    res = {
        "metric": METRIC, "value": entries / (ms / 1e3), "unit": "entries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int32", "data": "synthetic",
        "config": {"workload": (f"batched dense closure of {len(sizes)} pedigree components, sizes log-uniform "
                                f"in [16, 4096], layers = ceil(log2 n), p_parent=0.8, seed {args.seed}; one "
                                f"hipGraph replay per step (input reloaded by one device copy inside the step)"),
                   "components": len(sizes), "sum_n": int(sizes.sum()), "sum_n2": entries,
                   "closures_per_s": len(sizes) / (ms / 1e3),
                   "plain": "the same kernels launched one graph at a time on one stream",
                   "plain_ms": plain_ms,
                   "plain_closures_per_s": len(sizes) / (plain_ms / 1e3) if plain_ms else None,
                   "batched_vs_plain": plain_ms / ms if plain_ms else None,
                   "identical_to_plain": identical_plain, "graph_replay_identical_to_eager": identical_graph,
                   "single_matrix_path_ms": single_ms,
                   "batched_vs_single_matrix_path": single_ms / ms if single_ms else None,
                   "identical_to_single_matrix_path": identical_single,
                   "launches_per_batch": info["launches"], "pivot_steps": info["steps"],
                   "persistent_items": info["persistent_items"],
                   "tiles": {str(k): v for k, v in sorted(b.tiles.items())}, "schedule": b.schedule,
                   "autotune_s": tune_s},
        "roofline": roof,
    }
    res["config"]["persistent"] = _batched_persistent_leg(args, sizes, pristine, eager_result, b.stream)
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_batched(sizes, diag, edges)
    b.close()
    return res


def _batched_persistent_leg(args, sizes, pristine, eager_result, stream):
    """The same batch closed by the persistent launch (every component's work
    items in one grid): ms per graph replay, the launch's share and VALU
    fraction, and whether the pool equals the per-step launches' result."""
    from redblackgraph_amd import _lib
    from redblackgraph_amd.batched import BatchedClosure
    t = _lib.torch()
    L = _lib.load()
    p = BatchedClosure(sizes, tri=1, persistent=True, stream=stream)
    try:
        info = p.info()
        with t.cuda.stream(p.stream):
            p.pool.copy_(pristine)
        L.rbg_prof_enable(1)
        p.run(graph=False)
        p.stream.synchronize()
        launch_ms, _, relax = _prof_read(4)
        L.rbg_prof_enable(0)
        flags = sorted(set(int(s.flags) for s in p.statuses()))
        with t.cuda.stream(p.stream):
            for _ in range(args.warmup):
                p.pool.copy_(pristine)
                p.run(graph=True)
        p.stream.synchronize()
        t0 = time.perf_counter()
        with t.cuda.stream(p.stream):
            for _ in range(args.steps):
                p.pool.copy_(pristine)
                p.run(graph=True)
        p.stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        identical = bool(t.equal(p.pool, eager_result))
        achieved = relax * OPS_PER_RELAX["int32"] / (launch_ms / 1e3)
        return {"ms_per_step": ms, "launch_ms": launch_ms, "frac": achieved / PEAK_VALU_OPS,
                "relaxations_executed": relax, "work_items": info["persistent_items"],
                "launches_per_batch": info["launches"], "status_flags": flags,
                "identical_to_per_step_launches": identical,
                "kernel": "k_bfw_flow (every component's panel items and tiles in one persistent launch)"}
    finally:
        p.close()


def bench_serve(args):
    """Served closure stream: Poisson arrivals, dense int32 below 4096
    vertices and CSR int64 above, 1-64 edge insertions into resident graphs;
    value = closure entries produced per second of makespan."""
    import numpy as np
    from redblackgraph_amd import _lib, dense, synth
    from redblackgraph_amd._dtypes import elem_type
    from redblackgraph_amd.serving import ClosureServer, summarize, synthetic_trace, warmup_trace
    t = _lib.torch()
    trace = synthetic_trace(count=args.serve_count, rate=args.serve_rate, seed=args.seed)
    warm = synthetic_trace(count=12, rate=1000.0, seed=args.seed + 1000)
    srv = ClosureServer(hbm_budget=args.serve_budget_gib << 30, slots=args.serve_slots)
    try:
        srv.serve(warmup_trace(seed=args.seed))  # largest shapes: memory pools hold every block size
        srv.residents.clear()
        srv.used = 0
        srv.serve(warm)
        srv.residents.clear()
        srv.used = 0
        t.cuda.synchronize()
        t0 = time.perf_counter()
        res = srv.serve(trace)
        t.cuda.synchronize()
        span = time.perf_counter() - t0
    finally:
        srv.close()
    summ = summarize(res, span)
    # edge insertion roofline: one 64-edge insertion into a resident N=4096 closure
    n = 4096
    A = synth.dense_pedigree(synth.component_spec(n), 77)
    et = elem_type(np.int32)
    W, _ = dense.fw_closure(_lib.to_device(A), et)
    W = W.contiguous()
    rng = np.random.default_rng(1)
    diag = np.diagonal(A)
    edges = []
    for _ in range(64):
        a = int(rng.integers(n // 2, n))
        rel = int(rng.choice([2, 3]))
        cand = np.nonzero(diag[:a] == (-1 if rel == 2 else 1))[0]
        edges.append((a, int(rng.choice(cand)), rel))
    # the serving path: in-place compositions that touch the rows reaching the
    # edge's source; timed on the launch stream, then the same sequence again,
    # untimed, to count the bytes each composition must move
    fn = "rbg_edge_compose_inplace_i32"
    scratch = t.empty(int(_lib.load().rbg_edge_compose_scratch_bytes(n)), dtype=t.uint8, device=W.device)
    st = dense.status()
    sh = _lib.stream_handle(None)

    def compose(V, a, b, rel):
        _lib.call(fn, _lib.ptr(V), n, a, b, rel, _lib.ptr(scratch), et.width, 0, st.p, sh)

    V = W.clone()
    for a, b, rel in edges[:4]:
        compose(V, a, b, rel)
    V = W.clone()
    t.cuda.synchronize()
    ev0, ev1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    ev0.record()
    for a, b, rel in edges:
        compose(V, a, b, rel)
    ev1.record()
    ev1.synchronize()
    per_edge_s = ev0.elapsed_time(ev1) / 1e3 / len(edges)
    V2 = W.clone()
    b_total = 0
    for a, b, rel in edges:
        anc = int((V2[a] != 0).sum().item()) + (0 if int(V2[a, b].item()) != 0 else 1)
        aff = int((V2[:, a] != 0).sum().item()) - 1
        compose(V2, a, b, rel)
        supp = int((V2[a] != 0).sum().item())
        # composed row: its ancestors' rows read; column alpha read; affected
        # rows read and written where the composed row is present; row alpha written
        b_total += 4 * n * anc + 4 * n + 8 * supp * aff + 4 * n
    assert bool((V2 == V).all().item())
    b_edge = b_total / len(edges)
    roof = {"bound": "hbm", "kernel": "in-place edge insertion (k_edge_row + k_edge_update_rows), N=4096 int32",
            "achieved": b_edge / per_edge_s / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": b_edge / per_edge_s / PEAK_HBM, "traffic": None,
            "algorithmic_bytes_per_edge": b_edge, "ms_per_edge": per_edge_s * 1e3,
            "note": "seven small launches per edge: launch latency, not bytes, bounds an insertion into a pedigree "
                    "closure (the rows that reach an edge's source are few)",
            "bytes_rule": "4 N (ancestor rows of the source) + 4 N (source column) + 8 |composed row| (rows that "
                          "reach the source) + 4 N (composed row written)"}
    closes = [r for r in res if r.kind == "close" and r.error is None]
    return {
        "metric": METRIC, "value": summ["entries_per_s"], "unit": "entries/s",
        "n_gpus": 1, "steps": len(trace), "warmup": len(warm), "ms_per_step": span * 1e3 / max(1, len(trace)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32/int64",
        "data": "synthetic",
        "config": {"workload": (f"served closure stream: {len(trace)} requests, Poisson arrivals at "
                                f"{args.serve_rate}/s, new pedigrees (size log-uniform in [32, 16384]; dense int32 "
                                f"below 4096, CSR int64 above) or 1-64 parent-edge insertions into a resident "
                                f"closure (p=0.35); {args.serve_slots} slots, HBM budget {args.serve_budget_gib} GiB, "
                                f"seed {args.seed}"),
                   **{k: v for k, v in summ.items()},
                   "dense_closes": sum(1 for r in closes if r.n < 4096),
                   "sparse_closes": sum(1 for r in closes if r.n >= 4096),
                   "peak_reserved_gib": srv.peak / 2 ** 30, "evictions": srv.evictions},
        "roofline": roof,
    }


def cpu_baseline_order(Cs, perm_gpu, B_gpu=None):
    """The oracle's vectorised CSR ordering (numpy/scipy restatement of
    _ordering.pyx:50-157) and a scipy permutation, on the full closure.  Its
    permutation must equal the GPU's, and so must the permuted matrix (row
    pointers, column indices in row order, values) when the GPU's is given."""
    import numpy as np
    from oracle import oracle
    t0 = time.perf_counter()
    perm, _ = oracle.canonical_ordering_csr(Cs)
    B = Cs[perm][:, perm]
    B.sort_indices()
    dt = time.perf_counter() - t0
    assert np.array_equal(perm, perm_gpu), "GPU canonical order differs from the oracle's"
    if B_gpu is not None:
        ip, ix, dv = B_gpu
        assert np.array_equal(ip, B.indptr) and np.array_equal(ix, B.indices) and np.array_equal(dv, B.data), \
            "GPU permuted closure differs from the oracle's"
    return {
        "value": Cs.nnz / dt, "unit": "entries/s", "cores": 1, "kind": "port",
        "sample": (f"oracle CSR canonical ordering (numpy/scipy restatement of _ordering.pyx:50-157 and "
                   f"_permutation.pyx:67-136, 1 thread) on the full closure: {dt:.2f} s for {Cs.nnz} entries; "
                   f"host CPU: {_cpu_model()}"),
    }


def bench_order(args):
    """Canonical ordering (components, metrics, sort, CSR permutation) of the
    1M-vertex closure, on the device."""
    import numpy as np
    from redblackgraph_amd import _lib, synth
    from redblackgraph_amd import sparse_dev as sd
    from redblackgraph_amd.sparse.csgraph._ordering import canonical_order_device, _permute_csr_device
    t = _lib.torch()
    A = synth.csr_pedigree(synth.SPARSE_1M, args.seed)
    C, info = sd.dag_closure(sd.canonicalize(sd.from_scipy(A, np.int64)))
    n, nnz = C.n_rows, C.nnz
    for _ in range(max(1, args.warmup)):
        perm, B, counts = canonical_order_device(C)
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        perm, B, counts = canonical_order_device(C)
    t.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the permutation alone, timed with events on the stream it runs on
    dt = np.dtype(np.int64)
    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        _permute_csr_device(C.indptr, C.indices, C.data, n, nnz, dt, perm)
    e1.record()
    t.cuda.synchronize()
    perm_ms = e0.elapsed_time(e1) / args.steps
    # size-independent checks: a permutation of the vertices, the entry count
    # kept, and an upper triangular result (a closed pedigree in canonical order)
    p = perm.cpu().numpy()
    assert np.array_equal(np.sort(p), np.arange(n))
    rows = t.repeat_interleave(t.arange(n, device=B.indptr.device), B.indptr[1:] - B.indptr[:-1])
    below = int((B.indices.to(t.int64) < rows).sum().item())
    assert B.nnz == nnz and below == 0
    bytes_alg = 2 * (12 * nnz + 8 * (n + 1)) + 4 * n
    achieved = bytes_alg / (perm_ms / 1e3)
    roof = {
        "bound": "hbm", "kernel": "rbg_permute_csr<long> (rows renamed and copied, then sorted per row)",
        "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": achieved / PEAK_HBM,
        "traffic": _read_traffic("permute_csr_order1m_pmc.txt"), "algorithmic_bytes_per_launch": bytes_alg,
        "avg_launch_ms": perm_ms,
    }
    res = {
        "metric": "canonical ordering entries/sec (components, keys, sort, CSR permutation of a closure)",
        "value": nnz / (ms / 1e3), "unit": "entries/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "int64", "data": "synthetic",
        "config": {"workload": f"canonical ordering of the 1M-vertex pedigree closure (CSR, int64), seed {args.seed}",
                   "n": n, "nnz": nnz, "components": int(counts.numel())},
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        Bg = (B.indptr.cpu().numpy(), B.indices.cpu().numpy(), B.data.cpu().numpy())
        res["cpu_baseline"] = cpu_baseline_order(sd.to_scipy(C), p, Bg)
        res["config"]["identical_to_oracle"] = "vertex order and the permuted CSR (indptr, indices, data)"
    return res


def cpu_baseline_matmul(n_sample: int, n_target: int):
    """Single-core 'ij,jk' loop (the oracle's restatement of redblack.c.src:
    554-700) on the closure of an n_sample pedigree, N^3-scaled."""
    from oracle import oracle
    from redblackgraph_amd import synth
    A = synth.dense_pedigree(synth.PedigreeSpec(n_sample, max(4, round(24 * n_sample / 8192)), 0.75, 4), 1)
    W = oracle.warshall(A).W
    with _one_core() as pin:
        t0 = time.perf_counter()
        oracle.matmul(W, W)
        dt = time.perf_counter() - t0
    rate = n_sample ** 3 / dt
    return {"value": rate / n_target, "unit": "entries/s", "cores": 1, "kind": "port",
            "sample": (f"oracle 'ij,jk' product loop (C restatement of redblack.c.src:554-700, gcc -O2, 1 thread) "
                       f"on the closure of a seeded N={n_sample} pedigree, squared: {dt:.2f} s = {rate:.3e} terms/s; "
                       f"value = that rate / {n_target} (N^3 scaling); {pin.host()}")}


def bench_matmul(args):
    """Dense AVOS product W @ W of the N=8192 int32 closure (the rb.array @
    path, redblack.c.src:554-700): the fast kernel timed with events on its
    stream; checked against the general kernel at full size."""
    import numpy as np
    from redblackgraph_amd import _lib, dense, synth
    from redblackgraph_amd._dtypes import elem_type
    t = _lib.torch()
    n = 8192
    A = synth.dense_pedigree(synth.DENSE_8192, args.seed)
    et = elem_type(np.int32)
    W, _ = dense.fw_closure(_lib.to_device(A), et)
    W = W.contiguous()
    for _ in range(max(1, args.warmup)):
        C, s = dense.matmul(W, W, et)
    if s.flags & _lib.F_NEED_GENERAL:
        raise RuntimeError("the fast product fell back to the general kernel")
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        C, s = dense.matmul(W, W, et)
    t.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    # the product kernel alone, with events on its stream (the current stream)
    ws = t.empty(int(_lib.load().rbg_matmul_fast_workspace_bytes(n, n, n, 4)), dtype=t.uint8, device=W.device)
    st = dense.status()
    sh = _lib.stream_handle(None)
    e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(args.steps):
        _lib.call("rbg_matmul_fast_i32", _lib.ptr(W), n, _lib.ptr(W), n, _lib.ptr(C), n, n, n, n, _lib.ptr(ws),
                  ws.numel(), st.p, sh)
    e1.record()
    t.cuda.synchronize()
    kernel_ms = e0.elapsed_time(e1) / args.steps
    # full-size parity: the general kernel's product (full arithmetic)
    Cg, sg = dense.matmul(W, W, et, force_general=True)
    identical = bool(t.equal(C, Cg))
    if not identical or sg.events:
        raise RuntimeError("fast and general products differ")
    terms = float(n) ** 3
    # the product's inner loop has no trajectory maximum: per cell and two
    # steps two v_lshl_or_b32 and one v_min3_u32, 1.5 VALU ops per term
    achieved = terms * MM_OPS_PER_TERM / (kernel_ms / 1e3)
    res = {
        "metric": "dense AVOS product entries/sec (W @ W of a closure)", "value": n * n / (ms / 1e3),
        "unit": "entries/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": f"W @ W, W the closure of the seeded N=8192 int32 pedigree, seed {args.seed}",
                   "n": n, "identical_to_general_kernel": identical, "kernel_ms": kernel_ms},
        "roofline": {"bound": "valu", "kernel": "k_mm_fast<int, 8> (phase-3 tile over all k)",
                     "achieved": achieved / 1e12, "peak": PEAK_VALU_OPS / 1e12, "unit": "TOP/s",
                     "frac": achieved / PEAK_VALU_OPS, "traffic": _read_traffic("mm_fast_matmul8192_pmc.txt"),
                     "ops_per_term": MM_OPS_PER_TERM, "terms_per_launch": terms, "avg_launch_ms": kernel_ms},
    }
    if not args.no_cpu_baseline:
        res["cpu_baseline"] = cpu_baseline_matmul(1024, n)
    return res


def bench_hbmcopy(args):
    """A plain device-to-device copy of 1 GiB (reads 1 GiB, writes 1 GiB):
    the calibration run for the HBM byte counters (FETCH_SIZE, WRITE_SIZE)."""
    from redblackgraph_amd import _lib
    t = _lib.torch()
    nbytes = 1 << 30
    src = t.empty(nbytes // 4, dtype=t.int32, device=_lib.device()).fill_(7)
    dst = t.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    t.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    return {"metric": "device copy bandwidth (calibration)", "value": 2 * nbytes / (ms / 1e3) / 1e9,
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
            "data": "synthetic", "config": {"workload": "1 GiB int32 device-to-device copy", "bytes_read": nbytes,
                                            "bytes_written": nbytes}}


def _sub(res: dict) -> dict:
    """A workload's result without the keys the top-level line already holds."""
    drop = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data")
    return {k: v for k, v in res.items() if k not in drop}


def bench_all(args):
    """The whole metric in one line: the dense N=8192 closure is the top-level
    value; the sparse 1M-vertex closure, the batched 4096-component closure
    (against both per-graph paths), the N=4096 int64 closure, the served
    stream, the dense N=8192 product and the canonical ordering of the 1M
    closure are nested under sub_results, each with its own steps, roofline
    and CPU baseline."""
    import copy
    res = bench_dense(args, 8192, "int32")
    subs = {}
    a = copy.copy(args)
    print("bench: sparse1m", file=sys.stderr, flush=True)
    subs["sparse1m"] = _sub(bench_sparse(a))
    a = copy.copy(args)
    a.steps, a.warmup = max(2, args.steps // 5), 1
    print("bench: batched", file=sys.stderr, flush=True)
    subs["batched"] = _sub(bench_batched(a))
    a = copy.copy(args)
    print("bench: dense4096_i64", file=sys.stderr, flush=True)
    subs["dense4096_i64"] = _sub(bench_dense(a, 4096, "int64"))
    a = copy.copy(args)
    print("bench: serve", file=sys.stderr, flush=True)
    subs["serve"] = _sub(bench_serve(a))
    a = copy.copy(args)
    a.steps, a.warmup = max(2, args.steps // 2), 1
    print("bench: matmul8192", file=sys.stderr, flush=True)
    subs["matmul8192"] = _sub(bench_matmul(a))
    a = copy.copy(args)
    print("bench: order1m", file=sys.stderr, flush=True)
    subs["order1m"] = _sub(bench_order(a))
    res["sub_results"] = subs
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workload", default="all",
                    choices=["all", "dense8192", "dense4096_i64", "dense512", "sparse1m", "batched", "serve",
                             "order1m", "matmul8192", "hbmcopy"])
    ap.add_argument("--serve-count", type=int, default=300)
    ap.add_argument("--serve-rate", type=float, default=150.0)
    ap.add_argument("--serve-slots", type=int, default=4)
    ap.add_argument("--serve-budget-gib", type=int, default=64)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--cpu-sample", type=int, default=2048,
                    help="order of the dense CPU baseline's sample, run in full and N^3-scaled to the workload "
                         "(about 26 s at 2048 on one core)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prof-every", type=int, default=4,
                    help="bracket the kernels of every N-th timed step with timing events (1: every step)")
    ap.add_argument("--no-plain", action="store_true",
                    help="batched workload: skip the one-graph-at-a-time comparison legs (profiling runs)")
    ap.add_argument("--dense-schedule", default="auto", choices=["auto", "steps"],
                    help="dense workloads: 'steps' keeps one launch per pivot step for lower triangular "
                         "inputs instead of the persistent closure (a comparison run)")
    args = ap.parse_args()
    args.prof_every = max(1, args.prof_every)
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "all":
        res = bench_all(args)
    elif args.workload == "dense8192":
        res = bench_dense(args, 8192, "int32")
    elif args.workload == "sparse1m":
        res = bench_sparse(args)
    elif args.workload == "hbmcopy":
        res = bench_hbmcopy(args)
    elif args.workload == "matmul8192":
        res = bench_matmul(args)
    elif args.workload == "order1m":
        res = bench_order(args)
    elif args.workload == "serve":
        res = bench_serve(args)
    elif args.workload == "batched":
        res = bench_batched(args)
    elif args.workload == "dense4096_i64":
        res = bench_dense(args, 4096, "int64")
    else:
        res = bench_dense(args, 512, "int32")
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
