"""csgraph closure family (sparse/csgraph/__init__.py of the reference)."""
from .cycleerror import CycleError
from .transitive_closure import transitive_closure, transitive_closure_floyd_warshall

__all__ = ["CycleError", "transitive_closure", "transitive_closure_floyd_warshall"]
