"""Independent checkers used to pin the oracle (and, through it, the GPU path).

These do not restate the reference's loops: they compute the same quantities
from the genealogical meaning of the AVOS encoding.
"""
from __future__ import annotations

import numpy as np


def ahnentafel_closure(A: np.ndarray) -> np.ndarray:
    """Closure of a valid pedigree matrix by breadth-first search.

    For a parity-consistent pedigree (A[child, father] = 2 with a red father,
    A[child, mother] = 3 with a black mother, diagonal +-1) the closure entry
    (i, j) is the smallest Ahnentafel number of ancestor j seen from i: the
    root is 1, a father appends bit 0 (2a), a mother bit 1 (2a + 1).  The
    diagonal keeps its colour.  Built root-outwards, unlike the AVOS product,
    which composes from the far end.
    """
    n = A.shape[0]
    W = np.zeros((n, n), dtype=np.int64)
    parents = [[(j, int(A[i, j])) for j in range(n) if j != i and A[i, j] != 0] for i in range(n)]
    for i in range(n):
        W[i, i] = A[i, i]
        frontier = [(i, 1)]
        best = {}
        while frontier:
            nxt = []
            for v, num in frontier:
                for p, rel in parents[v]:
                    cand = 2 * num + (0 if rel == 2 else 1)
                    if p not in best or cand < best[p]:
                        best[p] = cand
                        nxt.append((p, cand))
            frontier = nxt
        for p, num in best.items():
            if p != i:
                W[i, p] = num
    return W


def first_discovery_diameter(A: np.ndarray) -> int:
    """MSB of the largest value the naive k-i-j loop ever writes, for a valid pedigree.

    A cell (i, j) first becomes nonzero at the smallest k such that some path
    i -> ... -> j uses only intermediates <= k (k = -1 for a direct edge); its
    value then is the smallest path number among those paths and it only
    decreases afterwards, so the trajectory maximum is that first value.
    Enumerates every path explicitly (small graphs only).
    """
    n = A.shape[0]
    parents = [[(j, int(A[i, j])) for j in range(n) if j != i and A[i, j] != 0] for i in range(n)]
    best = 0
    for i in range(n):
        if A[i, i] == 1:
            best = max(best, 1)
        paths = {}  # j -> list of (max_intermediate, number)
        stack = [(i, 1, -1)]
        while stack:
            v, num, mx = stack.pop()
            for p, rel in parents[v]:
                cand = 2 * num + (0 if rel == 2 else 1)
                paths.setdefault(p, []).append((mx, cand))
                stack.append((p, cand, max(mx, p)))
        for j, lst in paths.items():
            if j == i:
                continue
            kf = min(m for m, _ in lst)
            first = min(v for m, v in lst if m <= kf)
            best = max(best, first)
    return best.bit_length() - 1 if best > 1 else 0


def random_valid_pedigree(n: int, rng: np.random.Generator, p: float = 0.7, lookback: int = 4) -> np.ndarray:
    """Small random pedigree (lower triangular, parity consistent)."""
    red = rng.random(n) < 0.5
    A = np.zeros((n, n), dtype=np.int64)
    for i in range(n):
        A[i, i] = -1 if red[i] else 1
        lo = max(0, i - lookback)
        males = [j for j in range(lo, i) if red[j]]
        females = [j for j in range(lo, i) if not red[j]]
        if males and rng.random() < p:
            A[i, rng.choice(males)] = 2
        if females and rng.random() < p:
            A[i, rng.choice(females)] = 3
    return A


def dense_triple_loop(A, B, prod, add):
    n, m = A.shape
    _, p = B.shape
    C = np.zeros((n, p), dtype=np.int64)
    for i in range(n):
        for k in range(p):
            acc = 0
            for j in range(m):
                acc = add(acc, prod(int(A[i, j]), int(B[j, k])))
            C[i, k] = acc
    return C


def overflow_chain(n, dt):
    """A paternal chain of n vertices whose middle vertex has the largest
    index: the naive loop meets MSB(x) + MSB(y) = width only at the last step,
    with both operands still small, so the reference raises OverflowError
    (instead of first storing a wrapped negative it would later loop on)."""
    mid = (n - 1) // 2
    idx = lambda t: t if t < mid else (n - 1 if t == mid else t - 1)
    A = np.zeros((n, n), dtype=dt)
    np.fill_diagonal(A, -1)
    for t in range(n - 1):
        A[idx(t + 1), idx(t)] = 2
    return A
