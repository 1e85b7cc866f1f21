"""Cooperative CSR products (sparse_operator.py of the reference, on gfx950).

``SparseOperator(A, row_lanes)`` exposes the multi-lane CSR product as a linear
operator; ``_matvec`` is the shared product the factors use (reference
sparse_operator.py:90-104), here always the HIP kernel ``wpc_csr_spmv``.
"""

from __future__ import annotations

import torch

from . import _native as nat
from .linear import LinearOperator
from .sparse import BsrMatrix, StreamPlan, bsr_copy

LANES = (1, 2, 4, 8, 16, 32)


def _matvec(a: BsrMatrix, x, y, z, alpha, beta, lanes, row_scale=None, nvec=1):
    """z = alpha * A x + beta * y for canonical scalar CSR; x and z must not alias.

    Storage narrower than the vectors (float32/float16/bfloat16 factors with
    float64 vectors) accumulates in the vector type (sparse_operator.py:64-87).
    """
    nat.call("wpc_csr_spmv", a.nrow, nat.ptr(a.offsets), nat.ptr(a.columns), nat.ptr(a.values),
             nat.code(a.scalar_type), nat.ptr(row_scale), nat.ptr(x), nat.ptr(y), nat.ptr(z),
             nat.code(x.dtype), nvec, float(alpha), float(beta), int(lanes), nat.stream())


class SparseOperator(LinearOperator):
    """BSR linear operator with ``row_lanes`` lanes cooperating on each scalar row.

    Construction scalarizes on the device.  Owns scratch for x/z aliasing; do not
    apply one instance concurrently on independent streams (sparse_operator.py:107-139).
    """

    def __init__(self, A, row_lanes=4):
        if not isinstance(A, BsrMatrix):
            raise TypeError("SparseOperator requires a Warp BSR matrix")
        if A.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("SparseOperator supports float32 and float64")
        if not isinstance(row_lanes, int) or row_lanes not in LANES:
            raise ValueError("row_lanes must be one of 1, 2, 4, 8, 16, 32")
        nat.require_device(A.values, "A")
        self.source = A
        self.scalar = bsr_copy(A, block_shape=(1, 1)) if A.block_shape != (1, 1) else bsr_copy(A)
        self.row_lanes = row_lanes
        self._alias = torch.empty(self.scalar.ncol, dtype=A.scalar_type, device=A.device)
        self._plan = None
        super().__init__(A.shape, A.dtype, A.device, self._apply)

    fused = True

    def _stream(self):
        if self._plan is None:
            self._plan = StreamPlan(self.scalar, None if self.row_lanes == 1 else self.row_lanes)
        return self._plan

    def _apply(self, x, y, z, alpha, beta):
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        if x.data_ptr() == z.data_ptr():
            self._alias.copy_(x)
            x = self._alias
        self._stream().run(x, y, z, alpha, beta)

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        """z = alpha A x + beta y on nvec interleaved vectors (+ fused dot)."""
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        if x.data_ptr() == z.data_ptr():
            x = x.clone()
        return self._stream().run(x, y, z, alpha, beta, nvec,
                                  w=None if dot_w is None else dot_w.reshape(-1), out=dot_out)
