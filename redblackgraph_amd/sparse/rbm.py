"""rb_matrix: scipy CSR whose matrix product is the AVOS product
(sparse/rbm.py:9-122 of the reference), computed on the GPU.

The product follows rbm.py:39-122: operands are canonicalised, values are
promoted to 64 bits (the unsigned RED_ONE sentinel remapped), the product is
formed, zero sums are dropped and the result is cast back to the upcast
dtype.  Output rows come back sorted (the reference's pass 2 leaves them in
linked-list order; compare canonical forms)."""
from __future__ import annotations

import numpy as np
from scipy.sparse import csr_matrix

from .. import sparse_dev as sd


class rb_matrix(csr_matrix):

    def __matmul__(self, other):
        return self._mul_sparse_matrix(other)

    def __rmatmul__(self, other):
        return self.__class__(other)._mul_sparse_matrix(self)

    def __mul__(self, other):
        if np.isscalar(other):
            return super().__mul__(other)
        return self._mul_sparse_matrix(other)

    def __rmul__(self, other):
        if np.isscalar(other):
            return super().__rmul__(other)
        return self.__class__(other)._mul_sparse_matrix(self)

    def transitive_closure(self, method="D"):
        from .csgraph import transitive_closure
        return transitive_closure(self, method)

    def _mul_sparse_matrix(self, other):
        from scipy.sparse._sputils import upcast
        M, K1 = self.shape
        other = self.__class__(other)
        K2, N = other.shape
        if K1 != K2:
            raise ValueError(f"dimension mismatch: {self.shape} @ {other.shape}")
        out_dtype = np.dtype(upcast(self.dtype, other.dtype))
        wide = np.dtype(np.uint64) if out_dtype.kind == "u" else np.dtype(np.int64)

        def widen(m):
            d = np.asarray(m.data).astype(wide)
            if m.dtype.kind == "u" and m.dtype.itemsize < 8:
                d[np.asarray(m.data) == np.iinfo(m.dtype).max] = np.iinfo(wide).max
            return csr_matrix((d.view(np.int64), m.indices, m.indptr), shape=m.shape)

        a = sd.canonicalize(sd.from_scipy(widen(self), np.int64))
        b = sd.canonicalize(sd.from_scipy(widen(other), np.int64))
        C, events = sd.spgemm(a, b, width=64, is_unsigned=out_dtype.kind == "u")
        if events:
            raise OverflowError("Avos product results in an overflow.")
        c = sd.to_scipy(C)
        data = c.data.view(np.uint64) if out_dtype.kind == "u" else c.data
        return rb_matrix((data.astype(out_dtype, copy=False), c.indices, c.indptr), shape=(M, N))
