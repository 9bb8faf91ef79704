"""csgraph closure family (sparse/csgraph/__init__.py of the reference)."""
from .cycleerror import CycleError
from ._shortest_path import shortest_path, floyd_warshall
from .transitive_closure import (
    transitive_closure,
    transitive_closure_floyd_warshall,
    transitive_closure_dijkstra,
    transitive_closure_squaring,
    transitive_closure_adaptive,
    transitive_closure_dag_sparse,
    DensificationError,
)
from ._components import (
    component_wise_closure,
    find_components,
    find_components_sparse,
    extract_submatrix,
    merge_component_matrices,
    get_component_vertices,
)
from ._ordering import avos_canonical_ordering, permute, permute_sparse, _get_permutation, _get_permutation_sparse
from ._topological_sort import topological_sort, topological_ordering, is_upper_triangular
from ._relational_composition import vertex_relational_composition, edge_relational_composition

__all__ = ["CycleError", "DensificationError", "shortest_path", "floyd_warshall", "transitive_closure",
           "transitive_closure_floyd_warshall", "transitive_closure_dijkstra", "transitive_closure_squaring",
           "transitive_closure_adaptive", "transitive_closure_dag_sparse", "component_wise_closure",
           "find_components", "find_components_sparse", "extract_submatrix", "merge_component_matrices",
           "get_component_vertices", "avos_canonical_ordering", "permute", "permute_sparse",
           "topological_sort", "topological_ordering", "is_upper_triangular",
           "vertex_relational_composition", "edge_relational_composition"]
