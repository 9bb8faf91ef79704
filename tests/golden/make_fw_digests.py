"""Writes tests/golden/fw_digests.json: the oracle's naive Floyd-Warshall
(the C restatement of warshall.c.in:50-107) run once, in full, on the
full-size dense workloads, reduced to a digest of W and the trajectory
diameter so the GPU tests can compare the 8192-vertex closure without
running the oracle there.

    python tests/golden/make_fw_digests.py [name ...]

Names: dense8192_i32 (about 30 min on one core), dense4096_i64 (about
5 min), dense2304_i32 (under a minute).  Entries already in the file are
kept unless named again.  The digest is sha256 over W's bytes (C order,
little-endian, the workload's dtype).
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from redblackgraph_amd import synth  # noqa: E402

OUT = os.path.join(HERE, "fw_digests.json")

SPEC_2304 = synth.PedigreeSpec(2304, 16, 0.75, 4)

# name -> (spec, seed, dtype, renumbering): "" keeps the generator's order
# (lower triangular), "shuffle:<seed>" renumbers by synth.shuffle_vertices,
# "reverse" by synth.reverse_vertices (upper triangular)
WORKLOADS = {
    "dense8192_i32": (synth.DENSE_8192, 0, np.int32, ""),
    "dense4096_i64": (synth.DENSE_4096, 0, np.int64, ""),
    "dense2304_i32": (SPEC_2304, 0, np.int32, ""),
    "dense2304_shuffled_i32": (SPEC_2304, 0, np.int32, "shuffle:1"),
    "dense2304_reversed_i32": (SPEC_2304, 0, np.int32, "reverse"),
}


def workload_matrix(name: str) -> np.ndarray:
    spec, seed, dtype, renum = WORKLOADS[name]
    A = synth.dense_pedigree(spec, seed, dtype)
    if renum.startswith("shuffle:"):
        A = synth.shuffle_vertices(A, int(renum.split(":")[1]))
    elif renum == "reverse":
        A = synth.reverse_vertices(A)
    return A


def digest(W: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(W).astype(W.dtype.newbyteorder("<")).tobytes()).hexdigest()


def run(name: str) -> dict:
    spec, seed, dtype, renum = WORKLOADS[name]
    A = workload_matrix(name)
    t0 = time.perf_counter()
    r = orc.warshall(A)
    dt = time.perf_counter() - t0
    W = r.W
    return {
        "spec": {"n": spec.n, "layers": spec.layers, "p_parent": spec.p_parent,
                 "lookback": spec.lookback, "p_red": spec.p_red},
        "seed": seed,
        "renumbering": renum,
        "dtype": np.dtype(dtype).name,
        "sha256_W": digest(W),
        "diameter": int(r.diameter),
        "nnz_W": int(np.count_nonzero(W)),
        "max_W": int(W.max()),
        "sum_W": int(W.astype(np.int64).sum()),
        "n_cycle": int(r.report["n_cycle"]),
        "n_overflow": int(r.report["n_overflow"]),
        "oracle_seconds": round(dt, 1),
    }


def main(names):
    for name in names or list(WORKLOADS):
        print(f"oracle warshall: {name} ...", flush=True)
        entry = run(name)
        print(f"  {entry}", flush=True)
        data = {}
        if os.path.exists(OUT):  # read again: another run may have added entries meanwhile
            with open(OUT) as f:
                data = json.load(f)
        data[name] = entry
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
            f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:])
