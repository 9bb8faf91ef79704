"""GPU backend, shaped like redblackgraph.gpu (gpu/__init__.py:14-73 of the
reference): CSRMatrixGPU, spgemm, spgemm_upper_triangular, matmul_gpu,
transitive_closure_gpu, transitive_closure_dag_gpu and the element-wise
kernels -- backed by librbg_hip.so on MI355X instead of CuPy."""
from .avos_kernels import avos_sum_gpu, avos_product_gpu
from .csr_gpu import CSRMatrixGPU
from .spgemm import spgemm, spgemm_upper_triangular, matmul_gpu, spgemm_with_stats, SpGEMMStats
from .transitive_closure import transitive_closure_gpu, transitive_closure_dag_gpu

__all__ = [
    "CSRMatrixGPU",
    "spgemm", "spgemm_upper_triangular", "matmul_gpu",
    "transitive_closure_gpu", "transitive_closure_dag_gpu",
    "avos_sum_gpu", "avos_product_gpu",
]


def _available() -> bool:
    try:
        from .. import _lib
        _lib.load()
        return _lib.torch().cuda.is_available()
    except Exception:
        return False


# the reference exposes CUPY_AVAILABLE; the equivalent flag here
GPU_AVAILABLE = _available()
