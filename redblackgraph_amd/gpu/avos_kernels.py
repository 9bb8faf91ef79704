"""Element-wise AVOS on the GPU (gpu/avos_kernels.py:179-297 of the reference).

``avos_sum_gpu`` / ``avos_product_gpu`` accept host arrays (returned as host
arrays) or torch tensors on the device (returned as device tensors); int32
and int64 are supported.
"""
from __future__ import annotations

import numpy as np

from .. import _lib

__all__ = ["avos_sum_gpu", "avos_product_gpu"]


def _prep(x, y):
    t = _lib.torch()
    on_dev = isinstance(x, t.Tensor) and x.is_cuda
    if not on_dev:
        x = np.asarray(x)
        y = np.asarray(y)
        if x.shape != y.shape:
            raise ValueError(f"Shape mismatch: {x.shape} vs {y.shape}")
        dt = np.int64 if (x.dtype == np.int64 or y.dtype == np.int64) else np.int32
        xd = _lib.to_device(x.astype(dt).reshape(-1))
        yd = _lib.to_device(y.astype(dt).reshape(-1))
        return xd, yd, x.shape, False
    if tuple(x.shape) != tuple(y.shape):
        raise ValueError(f"Shape mismatch: {tuple(x.shape)} vs {tuple(y.shape)}")
    dt = t.int64 if (x.dtype == t.int64 or y.dtype == t.int64) else t.int32
    return x.to(dt).contiguous().reshape(-1), y.to(dt).contiguous().reshape(-1), tuple(x.shape), True


def avos_sum_gpu(x, y):
    t = _lib.torch()
    xd, yd, shape, on_dev = _prep(x, y)
    z = t.empty_like(xd)
    fn = "rbg_avos_sum_i32" if xd.dtype == t.int32 else "rbg_avos_sum_i64"
    _lib.call(fn, _lib.ptr(xd), _lib.ptr(yd), _lib.ptr(z), xd.numel(), _lib.stream_handle())
    z = z.reshape(shape)
    return z if on_dev else z.cpu().numpy()


def avos_product_gpu(x, y):
    t = _lib.torch()
    xd, yd, shape, on_dev = _prep(x, y)
    z = t.empty_like(xd)
    fn = "rbg_avos_product_i32" if xd.dtype == t.int32 else "rbg_avos_product_i64"
    _lib.call(fn, _lib.ptr(xd), _lib.ptr(yd), _lib.ptr(z), xd.numel(), None, _lib.stream_handle())
    z = z.reshape(shape)
    return z if on_dev else z.cpu().numpy()
