"""Sparse AVOS API, shaped like redblackgraph.sparse (sparse/__init__.py,
rbm.py, csgraph/)."""
from ..core import avos_sum, avos_product
from .rbm import rb_matrix
from .csgraph import *  # noqa: F401,F403
from .csgraph import __all__ as _cs_all

__all__ = ["rb_matrix", "avos_sum", "avos_product"] + list(_cs_all)
