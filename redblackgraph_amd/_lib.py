"""ctypes binding of librbg_hip.so and the device plumbing around it.

The library is the product: if it is missing, or no GPU is present, every
entry point raises -- there is no CPU fallback anywhere in this package.
PyTorch supplies device memory and streams only.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librbg_hip.so")

_lib = None
_lock = threading.Lock()


class RbgStatus(ctypes.Structure):
    _fields_ = [
        ("events", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("cycle_count", ctypes.c_uint64),
        ("cycle_first_key", ctypes.c_uint64), ("cycle_last_key", ctypes.c_uint64),
        ("ov_first_key", ctypes.c_uint64), ("ov_last_key", ctypes.c_uint64),
        ("cycle_first_a", ctypes.c_int64), ("cycle_first_b", ctypes.c_int64),
        ("cycle_last_a", ctypes.c_int64), ("cycle_last_b", ctypes.c_int64),
        ("ov_last_a", ctypes.c_int64), ("ov_last_b", ctypes.c_int64),
        ("diameter_max", ctypes.c_int64),
        ("aux", ctypes.c_int64 * 8),
    ]


STATUS_BYTES = ctypes.sizeof(RbgStatus)

EV_OVERFLOW_WIDTH, EV_OVERFLOW_RED, EV_INVALID, EV_SIGN = 1, 2, 4, 8
F_NEED_GENERAL, F_PRECOND_FAIL, F_CYCLE, F_CAPACITY = 1, 2, 4, 8
F_HAS_LOWER, F_HAS_UPPER = 16, 32
F_STALLED = 64
NO_KEY = (1 << 64) - 1

# (name, restype, argtypes) of every exported symbol; include/rbg_hip.h is the
# authoritative declaration and tests/test_capi.py checks the two agree.
P, I32, I64, U64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64
SIGNATURES = {
    "rbg_status_init": (I32, [P, P]),
    "rbg_avos_sum_i32": (I32, [P, P, P, I64, P]),
    "rbg_avos_sum_i64": (I32, [P, P, P, I64, P]),
    "rbg_avos_product_i32": (I32, [P, P, P, I64, P, P]),
    "rbg_avos_product_i64": (I32, [P, P, P, I64, P, P]),
    "rbg_host_avos_sum_i64": (I64, [I64, I64]),
    "rbg_host_avos_product_u64": (U64, [I64, I64, ctypes.POINTER(ctypes.c_uint32)]),
    "rbg_fw_workspace_bytes": (I64, [I64, I32]),
    "rbg_fw_tri_workspace_bytes": (I64, [I64, I32]),
    "rbg_fw_flow_schedule": (I64, [I64, I32, I32, P, I64]),
    "rbg_fw_flow_wait_limit": (I64, [I64]),
    "rbg_fw_flow_waits": (I64, [I64, I32, I32, P, I64]),
    "rbg_fw_prof_enable": (I32, [I32]),
    "rbg_prof_enable": (I32, [I32]),
    "rbg_prof_read": (I32, [I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "rbg_prof_read_work": (I32, [I32, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                                 ctypes.POINTER(ctypes.c_double)]),
    "rbg_fw_prof_read": (I32, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                               ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "rbg_fw_precheck_i32": (I32, [P, I64, I64, P, P]),
    "rbg_fw_precheck_i64": (I32, [P, I64, I64, P, P]),
    "rbg_fw_load_i32": (I32, [P, I64, I64, P, I64, P, P]),
    "rbg_fw_load_i64": (I32, [P, I64, I64, P, I64, P, P]),
    "rbg_fw_dense_i32": (I32, [P, I64, I64, I32, I32, I32, P, I64, P, P]),
    "rbg_fw_dense_i64": (I32, [P, I64, I64, I32, I32, I32, P, I64, P, P]),
    "rbg_fw_dense_tri_i32": (I32, [P, I64, I64, I32, I32, I32, I32, P, I64, P, P]),
    "rbg_fw_dense_tri_i64": (I32, [P, I64, I64, I32, I32, I32, I32, P, I64, P, P]),
    "rbg_fw_noalloc_max_bytes": (I64, []),
    "rbg_fw_dense_noalloc_i32": (I32, [P, I64, I64, I32, I32, I32, P, P]),
    "rbg_fw_dense_noalloc_i64": (I32, [P, I64, I64, I32, I32, I32, P, P]),
    "rbg_matmul_i32": (I32, [P, I64, P, I64, P, I64, I64, I64, I64, I32, I32, P, P]),
    "rbg_matmul_i64": (I32, [P, I64, P, I64, P, I64, I64, I64, I64, I32, I32, P, P]),
    "rbg_matmul_fast_workspace_bytes": (I64, [I64, I64, I64, I32]),
    "rbg_matmul_precheck_i32": (I32, [P, I64, I64, I64, P, P]),
    "rbg_matmul_precheck_i64": (I32, [P, I64, I64, I64, P, P]),
    "rbg_matmul_fast_i32": (I32, [P, I64, P, I64, P, I64, I64, I64, I64, P, I64, P, P]),
    "rbg_matmul_fast_i64": (I32, [P, I64, P, I64, P, I64, I64, I64, I64, P, I64, P, P]),
    "rbg_edge_compose_i32": (I32, [P, I64, I64, I64, I32, P, I32, I32, P, P]),
    "rbg_edge_compose_i64": (I32, [P, I64, I64, I64, I64, P, I32, I32, P, P]),
    "rbg_edge_compose_scratch_bytes": (I64, [I64]),
    "rbg_edge_compose_inplace_i32": (I32, [P, I64, I64, I64, I32, P, I32, I32, P, P]),
    "rbg_edge_compose_inplace_i64": (I32, [P, I64, I64, I64, I64, P, I32, I32, P, P]),
    "rbg_vertex_compose_i32": (I32, [P, P, P, I32, I64, P, I32, I32, P, P]),
    "rbg_vertex_compose_i64": (I32, [P, P, P, I64, I64, P, I32, I32, P, P]),
    "rbg_csr_tmp_bytes": (I64, [I32]),
    "rbg_csr_canonicalize_i32": (I32, [P, P, P, I32, P, P, P, P, P]),
    "rbg_csr_canonicalize_i64": (I32, [P, P, P, I32, P, P, P, P, P]),
    "rbg_csr_validate": (I32, [P, P, I32, I32, P, P]),
    "rbg_csr_check_canonical_i32": (I32, [P, P, P, I32, P, P]),
    "rbg_csr_check_canonical_i64": (I32, [P, P, P, I32, P, P]),
    "rbg_dag_workspace_bytes": (I64, [I32, I64]),
    "rbg_dag_level_kinds": (I32, [P, I32, P, I32, P]),
    "rbg_dag_closure_i32": (I32, [P, P, P, I32, P, I64, P, P, I64, I32, P, I32, P, P]),
    "rbg_dag_closure_i64": (I32, [P, P, P, I32, P, I64, P, P, I64, I32, P, I32, P, P]),
    "rbg_topological_sort": (I32, [P, P, I32, P, I64, P, P, P, P]),
    "rbg_csr_triangles": (I32, [P, P, I32, P, P]),
    "rbg_dag_gather_i32": (I32, [P, I32, P, P, P, P, P, I64, P, P]),
    "rbg_dag_gather_i64": (I32, [P, I32, P, P, P, P, P, I64, P, P]),
    "rbg_spgemm_rows_i32": (I32, [P, P, P, I32, P, P, P, I32, P, P, P, I64, P, P, I64, I32, I32, P, P]),
    "rbg_csr_add_i32": (I32, [P, P, P, P, P, P, I32, I32, P, P, P, P, P, P, P, P]),
    "rbg_csr_add_i64": (I32, [P, P, P, P, P, P, I32, I32, P, P, P, P, P, P, P, P]),
    "rbg_spgemm_rows_i64": (I32, [P, P, P, I32, P, P, P, I32, P, P, P, I64, P, P, I64, I32, I32, P, P]),
    "rbg_batch_snap_bytes": (I64, [I32]),
    "rbg_batch_tri_snap_bytes": (I64, [I32]),
    "rbg_batch_flow_tickets": (I64, [P, P, I32, P, I64]),
    "rbg_batch_create": (I32, [P, I32, P]),
    "rbg_batch_info": (I32, [P, P]),
    "rbg_batch_run": (I32, [P, P]),
    "rbg_batch_capture": (I32, [P, P]),
    "rbg_batch_replay": (I32, [P, P]),
    "rbg_batch_destroy": (I32, [P]),
    "rbg_components": (I32, [P, P, I32, P, P, ctypes.POINTER(ctypes.c_int32), P]),
    "rbg_glue_temp_bytes": (I64, [I64]),
    "rbg_csr_expand_size": (I32, [P, P, I32, P, P, P]),
    "rbg_csr_max_abs_i32": (I32, [P, P, I32, P, P]),
    "rbg_csr_max_abs_i64": (I32, [P, P, I32, P, P]),
    "rbg_csr_to_dense_i32": (I32, [P, P, P, I32, I64, P, P]),
    "rbg_csr_to_dense_i64": (I32, [P, P, P, I32, I64, P, P]),
    "rbg_dense_to_csr_i32": (I32, [P, I32, I64, P, P, P, P, I64, P]),
    "rbg_dense_to_csr_i64": (I32, [P, I32, I64, P, P, P, P, I64, P]),
    "rbg_csr_column_i32": (I32, [P, P, P, I32, I32, P, P, P, P, I64, P]),
    "rbg_csr_column_i64": (I32, [P, P, P, I32, I32, P, P, P, P, I64, P]),
    "rbg_csr_row_set_i32": (I32, [P, P, P, I64, I32, I32, P, P, P, P]),
    "rbg_csr_row_set_i64": (I32, [P, P, P, I64, I32, I64, P, P, P, P]),
    "rbg_csr_upper_i32": (I32, [P, P, P, I32, P, P, P, P, I64, P]),
    "rbg_csr_upper_i64": (I32, [P, P, P, I32, P, P, P, P, I64, P]),
    "rbg_components_layout": (I32, [P, I32, P, P, P, P, P, P, P, P, I64, P]),
    "rbg_components_scatter": (I32, [P, P, P, I32, P, P, P, P, P, P]),
    "rbg_components_merge": (I32, [P, P, P, P, P, P, P, P, I32, P, P, P, P, I64, P]),
    "rbg_order_metrics_dense_i32": (I32, [P, I64, P, P, P, P]),
    "rbg_order_metrics_dense_i64": (I32, [P, I64, P, P, P, P]),
    "rbg_order_metrics_csr_i32": (I32, [P, P, P, I32, P, P, P, P]),
    "rbg_order_metrics_csr_i64": (I32, [P, P, P, I32, P, P, P, P]),
    "rbg_order_sort_temp_bytes": (I64, [I32]),
    "rbg_order_sort": (I32, [I32, P, P, P, P, P, P, P, I64, P]),
    "rbg_permute_dense_i32": (I32, [P, I64, P, P, I32, P]),
    "rbg_permute_dense_i64": (I32, [P, I64, P, P, I32, P]),
    "rbg_permute_csr_workspace_bytes_i32": (I64, [I32, I64]),
    "rbg_permute_csr_workspace_bytes_i64": (I64, [I32, I64]),
    "rbg_permute_csr_i32": (I32, [P, P, P, I32, I64, P, P, P, P, P, I64, P]),
    "rbg_permute_csr_i64": (I32, [P, P, P, I32, I64, P, P, P, P, P, I64, P]),
}


class RbgBatchComp(ctypes.Structure):
    """rbg_batch_comp: one component of a batched dense closure."""
    _fields_ = [("W", ctypes.c_void_p), ("snap", ctypes.c_void_p), ("st", ctypes.c_void_p),
                ("npad", ctypes.c_int32), ("tile", ctypes.c_int32), ("tri", ctypes.c_int32),
                ("persistent", ctypes.c_int32), ("snap_bytes", ctypes.c_int64)]


class LibraryMissing(ImportError):
    pass


def load():
    """Load librbg_hip.so (raises LibraryMissing if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise LibraryMissing(
                    f"{LIB_PATH} is not built; run `python -m redblackgraph_amd.build` "
                    "(the AVOS path has no CPU fallback)")
            # torch first, so that the library binds to the HIP runtime torch
            # has loaded rather than loading a second one
            torch()
            L = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def call(name, *args):
    rc = getattr(load(), name)(*args)
    if rc != 0:
        if rc == -1:
            raise ValueError(f"{name}: invalid argument")
        if rc == -2:
            raise ValueError(f"{name}: workspace too small")
        raise RuntimeError(f"{name}: HIP error {rc}")
    return rc


# ---------------------------------------------------------------------------
# torch plumbing
# ---------------------------------------------------------------------------
_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as _t
        _torch = _t
    return _torch


def device():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("redblackgraph_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "the AVOS path has no CPU fallback")
    return t.device("cuda", t.cuda.current_device())


def stream_handle(stream=None):
    t = torch()
    s = stream if stream is not None else t.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Status:
    """A device-resident rbg_status block plus a pinned host mirror."""

    def __init__(self):
        t = torch()
        self.dev = t.empty(STATUS_BYTES, dtype=t.uint8, device=device())
        self.host = t.empty(STATUS_BYTES, dtype=t.uint8, pin_memory=True)

    @property
    def p(self):
        return ptr(self.dev)

    def reset(self, stream=None):
        call("rbg_status_init", self.p, stream_handle(stream))

    def read(self, stream=None) -> RbgStatus:
        """Copy to host and synchronise the stream."""
        t = torch()
        s = stream if stream is not None else t.cuda.current_stream()
        with t.cuda.stream(s):
            self.host.copy_(self.dev, non_blocking=True)
        s.synchronize()
        return RbgStatus.from_buffer_copy(self.host.numpy().tobytes())


_TORCH_DT = {np.dtype(np.int32): "int32", np.dtype(np.int64): "int64"}


def torch_dtype(np_dtype):
    return getattr(torch(), _TORCH_DT[np.dtype(np_dtype)])


def to_device(a: np.ndarray):
    t = torch()
    return t.from_numpy(np.ascontiguousarray(a)).to(device(), non_blocking=False)
