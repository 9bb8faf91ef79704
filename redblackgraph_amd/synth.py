"""Seeded synthetic pedigrees for the benchmark and parity workloads.

Matrix convention of the reference synthesizer (util/synthesizer.py:481-509):
vertices are numbered layer by layer, founders first; A[child, father] = 2,
A[child, mother] = 3; the diagonal is RED_ONE (-1, male) or BLACK_ONE (1,
female).  Parents always precede children, so A is lower triangular and
acyclic.

Workload specs (sizes, layer counts, parent probabilities, look-back depth):
  dense N=512    10 layers, father/mother each known with p=0.8, from the previous 1-3 layers
  dense N=8192   24 layers, p=0.75, previous 1-4 layers (also N=4096 for int64)
  sparse 1M      15 layers, p=0.7, parents from the previous layer only
  batched        4096 components, sizes log-uniform in [16, 4096], layers = ceil(log2(size)), p=0.8
Generation uses numpy's PCG64 so a (spec, seed) pair names one graph.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class PedigreeSpec:
    n: int
    layers: int
    p_parent: float
    lookback: int
    p_red: float = 0.5


DENSE_512 = PedigreeSpec(512, 10, 0.8, 3)
DENSE_8192 = PedigreeSpec(8192, 24, 0.75, 4)
DENSE_4096 = PedigreeSpec(4096, 24, 0.75, 4)
SPARSE_1M = PedigreeSpec(1_000_005, 15, 0.7, 1)


def _layer_bounds(n: int, layers: int) -> np.ndarray:
    layers = max(1, min(layers, n))
    return np.linspace(0, n, layers + 1).round().astype(np.int64)


def pedigree_edges(spec: PedigreeSpec, seed: int):
    """Returns (is_red[n], child[], parent[], value[]) with value 2 (father) or 3 (mother)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = spec.n
    bounds = _layer_bounds(n, spec.layers)
    is_red = rng.random(n) < spec.p_red
    childs, parents, vals = [], [], []
    L = len(bounds) - 1
    for layer in range(1, L):
        lo, hi = bounds[layer], bounds[layer + 1]
        kids = np.arange(lo, hi, dtype=np.int64)
        if kids.size == 0:
            continue
        first = max(0, layer - spec.lookback)
        for sex_red, value in ((True, 2), (False, 3)):
            known = rng.random(kids.size) < spec.p_parent
            src_layer = rng.integers(first, layer, size=kids.size)
            # candidate parents of the right sex, per source layer
            pick = np.full(kids.size, -1, dtype=np.int64)
            for sl in range(first, layer):
                cand = np.arange(bounds[sl], bounds[sl + 1], dtype=np.int64)
                cand = cand[is_red[cand] == sex_red]
                sel = np.nonzero(src_layer == sl)[0]
                if cand.size == 0 or sel.size == 0:
                    continue
                pick[sel] = cand[rng.integers(0, cand.size, size=sel.size)]
            ok = known & (pick >= 0)
            childs.append(kids[ok])
            parents.append(pick[ok])
            vals.append(np.full(int(ok.sum()), value, dtype=np.int64))
    child = np.concatenate(childs) if childs else np.zeros(0, np.int64)
    parent = np.concatenate(parents) if parents else np.zeros(0, np.int64)
    value = np.concatenate(vals) if vals else np.zeros(0, np.int64)
    return is_red, child, parent, value


def dense_pedigree(spec: PedigreeSpec, seed: int, dtype=np.int32) -> np.ndarray:
    is_red, child, parent, value = pedigree_edges(spec, seed)
    A = np.zeros((spec.n, spec.n), dtype=dtype)
    A[child, parent] = value.astype(dtype)
    d = np.where(is_red, -1, 1).astype(dtype)
    A[np.arange(spec.n), np.arange(spec.n)] = d
    return A


def csr_pedigree(spec: PedigreeSpec, seed: int, dtype=np.int64):
    """CSR (indptr int32/int64, indices int32, data dtype) of the same graph."""
    import scipy.sparse as sp
    is_red, child, parent, value = pedigree_edges(spec, seed)
    n = spec.n
    rows = np.concatenate([np.arange(n, dtype=np.int64), child])
    cols = np.concatenate([np.arange(n, dtype=np.int64), parent])
    data = np.concatenate([np.where(is_red, -1, 1).astype(np.int64), value])
    A = sp.csr_matrix((data.astype(dtype), (rows, cols)), shape=(n, n))
    A.sum_duplicates()
    A.sort_indices()
    return A


def batched_sizes(count: int, lo: int, hi: int, seed: int) -> np.ndarray:
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(rng.uniform(math.log(lo), math.log(hi), size=count)).astype(np.int64).clip(lo, hi)


def component_spec(size: int, p_parent: float = 0.8, lookback: int = 3) -> PedigreeSpec:
    return PedigreeSpec(int(size), max(1, math.ceil(math.log2(max(2, size)))), p_parent, lookback)


def component_seed(seed: int, i: int) -> int:
    return seed * 1_000_003 + i


def batched_components(count: int = 4096, lo: int = 16, hi: int = 4096, seed: int = 0, dtype=np.int32):
    """List of dense component matrices for the batched workload."""
    sizes = batched_sizes(count, lo, hi, seed)
    return [dense_pedigree(component_spec(s), component_seed(seed, i), dtype) for i, s in enumerate(sizes)]


def component_edges(spec: PedigreeSpec, seed: int):
    """(diagonal[n] of -1/1, (child, parent, value)) of one pedigree: the same
    graph dense_pedigree(spec, seed) builds, without the n x n matrix."""
    is_red, child, parent, value = pedigree_edges(spec, seed)
    return np.where(is_red, -1, 1).astype(np.int32), (child, parent, value)


def batched_workload(count: int = 4096, lo: int = 16, hi: int = 4096, seed: int = 0):
    """Sizes, diagonals and edge lists of the batched workload (4096
    components, sizes log-uniform in [16, 4096], layers = ceil(log2(size)),
    each parent known with p=0.8 from the previous 1-3 layers)."""
    sizes = batched_sizes(count, lo, hi, seed)
    diag, edges = [], []
    for i, s in enumerate(sizes):
        d, e = component_edges(component_spec(int(s)), component_seed(seed, i))
        diag.append(d)
        edges.append(e)
    return sizes, diag, edges


def shuffle_vertices(A: np.ndarray, seed: int) -> np.ndarray:
    """The same graph with its vertices renumbered by a seeded permutation p
    (PCG64): B[i][j] = A[p[i]][p[j]].  A pedigree so renumbered is neither
    lower nor upper triangular."""
    p = np.random.Generator(np.random.PCG64(seed)).permutation(A.shape[0])
    return np.ascontiguousarray(A[np.ix_(p, p)])


def reverse_vertices(A: np.ndarray) -> np.ndarray:
    """The same graph numbered backwards (vertex i becomes n - 1 - i): a lower
    triangular pedigree becomes upper triangular."""
    return np.ascontiguousarray(A[::-1, ::-1])


def random_upper_dag(n: int, seed: int, hi: int = 4) -> np.ndarray:
    """The reference's random DAG construction (tests/sparse/test_dag_sparse_closure.py:77-83):
    legacy RandomState(seed), upper triangle of randint(0, hi), random +-1 diagonal."""
    rs = np.random.RandomState(seed)
    A = np.triu(rs.randint(0, hi, size=(n, n)), k=1)
    np.fill_diagonal(A, rs.choice([-1, 1], size=n))
    return A.astype(np.int32)
