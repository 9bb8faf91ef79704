"""redblackgraph_amd -- the AVOS closure path of redblackgraph on AMD MI355X.

Drop-in layout of the reference package's hot path:
  redblackgraph_amd.core            dense ``array`` with AVOS ``@`` and Floyd-Warshall closure
  redblackgraph_amd.sparse          ``rb_matrix`` and the csgraph closure family
  redblackgraph_amd.gpu             ``CSRMatrixGPU``, ``spgemm``, ``transitive_closure_gpu``,
                                    ``transitive_closure_dag_gpu``, element-wise kernels
  redblackgraph_amd.batched         many independent components closed in one graph launch
  redblackgraph_amd.serving         continuous-batching closure service
All arithmetic runs in librbg_hip.so (HIP, gfx950); nothing falls back to the CPU.
"""
__version__ = "0.1.0"

from .core import *  # noqa: F401,F403
from .core import __all__ as _core_all
from .constants import RED_ONE, BLACK_ONE, red_one_for_dtype, black_one_for_dtype, is_red_one, is_black_one
from .types import TransitiveClosure

__all__ = ["__version__", "TransitiveClosure", "RED_ONE", "BLACK_ONE", "red_one_for_dtype",
           "black_one_for_dtype", "is_red_one", "is_black_one"] + list(_core_all)


def __getattr__(name):
    # the sparse and gpu layers import scipy; load them on first use
    if name in ("sparse", "gpu", "batched", "serving", "synth"):
        import importlib
        return importlib.import_module(f"{__name__}.{name}")
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")
