"""Host-side component utilities (scipy in, scipy out): submatrix
extraction, merging and vertex grouping, against direct numpy indexing."""
import numpy as np
import scipy.sparse as sp

from redblackgraph_amd.sparse.csgraph import extract_submatrix, get_component_vertices, merge_component_matrices


def test_extract_and_merge_round_trip():
    rng = np.random.default_rng(0)
    n = 60
    D = (rng.random((n, n)) < 0.08) * rng.integers(2, 9, (n, n))
    np.fill_diagonal(D, 1)
    A = sp.csr_matrix(D.astype(np.int32))
    labels = (np.arange(n) % 3).astype(np.uint32)
    comps = get_component_vertices(labels)
    assert sorted(comps) == [0, 1, 2]
    parts = []
    for cid, verts in comps.items():
        sub, mapping = extract_submatrix(A, verts)
        np.testing.assert_array_equal(sub.toarray(), D[np.ix_(verts, verts)])
        np.testing.assert_array_equal(mapping, verts)
        parts.append((sub, mapping))
    merged = merge_component_matrices(parts, n)
    expect = np.zeros_like(D)
    for verts in comps.values():
        expect[np.ix_(verts, verts)] = D[np.ix_(verts, verts)]
    np.testing.assert_array_equal(merged.toarray(), expect)
    assert merged.has_sorted_indices


def test_get_component_vertices_sorted_groups():
    labels = np.array([2, 0, 2, 1, 0, 2], dtype=np.uint32)
    g = get_component_vertices(labels)
    assert {k: v.tolist() for k, v in g.items()} == {0: [1, 4], 1: [3], 2: [0, 2, 5]}
