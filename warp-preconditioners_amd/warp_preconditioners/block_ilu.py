"""Block Jacobi and synchronous block ILU(0) on gfx950 (drop-in for block_ilu.py).

Square BSR blocks of size 1..4 in float32/float64.  Setup synchronizes to read
validation counters; apply is stream-ordered and capturable.  Instances own
scratch; do not apply one instance on concurrent streams.
"""

from __future__ import annotations

import math

import torch

from . import _native as nat
from .linear import LinearOperator
from .sparse import BsrMatrix, bsr_copy


class BlockJacobi(LinearOperator):
    """Invert diagonal scalar, 2x2, 3x3 or 4x4 BSR blocks on device (block_ilu.py:225-344).

    Blocks may be indefinite but must be nonsingular: a singular or near-singular
    block (max-abs-scaled determinant at most ``pivot_tolerance``) raises.
    """

    def __init__(self, A, pivot_tolerance=None):
        if not isinstance(A, BsrMatrix) or A.nrow != A.ncol:
            raise ValueError("A must be a square Warp BSR matrix")
        size = A.block_shape[0]
        if A.block_shape != (size, size) or size not in (1, 2, 3, 4):
            raise ValueError("BlockJacobi requires square blocks of size 1 through 4")
        if A.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("A must use float32 or float64")
        tolerance = ((1e-12 if A.scalar_type == torch.float64 else 1e-6)
                     if pivot_tolerance is None else pivot_tolerance)
        if not math.isfinite(tolerance) or tolerance < 0:
            raise ValueError("pivot_tolerance must be finite and nonnegative")
        nat.require_device(A.values, "A")
        self.size = size
        self.pivot_tolerance = tolerance
        self._bad = torch.zeros(1, dtype=torch.int32, device=A.device)
        shape = (A.nrow, size, size)
        if type(self) is BlockJacobi:
            self.A = A
            self.inverse_diagonal = torch.empty(shape, dtype=A.scalar_type, device=A.device)
            self._candidate = torch.empty_like(self.inverse_diagonal)
            super().__init__(A.shape, A.dtype, A.device, self._apply)
            self.update(A)
            return
        self.A = bsr_copy(A)
        self._rows = torch.empty(self.A.nnz, dtype=torch.int32, device=A.device)
        self._diagonal = torch.empty(A.nrow, dtype=torch.int32, device=A.device)
        self.inverse_diagonal = torch.zeros(shape, dtype=A.scalar_type, device=A.device)
        nat.call("wpc_block_topology", A.nrow, nat.ptr(self.A.offsets), nat.ptr(self.A.columns),
                 nat.ptr(self._rows), nat.ptr(self._diagonal), nat.ptr(self._bad), nat.stream())
        self._invert(self.A.values)
        self._check()
        super().__init__(A.shape, A.dtype, A.device, self._apply)

    def update(self, A=None):
        """Recompute diagonal inverses into existing buffers; a failed update changes
        nothing (block_ilu.py:282-317)."""
        if type(self) is not BlockJacobi:
            raise NotImplementedError("Rebuild BlockILU0 to update its factors")
        if A is None:
            A = self.A
        if not isinstance(A, BsrMatrix) or (A.shape, A.dtype, A.device) != (
            self.shape, self.dtype, self.device,
        ):
            raise ValueError("update requires the same shape, dtype, and device")
        self._bad.zero_()
        nat.call("wpc_block_diagonal_only", A.nrow, self.size, nat.ptr(A.offsets),
                 nat.ptr(A.columns), nat.ptr(A.values), nat.ptr(self._candidate),
                 float(self.pivot_tolerance), nat.ptr(self._bad), nat.code(A.scalar_type),
                 nat.stream())
        self._check()
        self.inverse_diagonal.copy_(self._candidate)
        self.A = A
        return self

    def _invert(self, values):
        nat.call("wpc_block_invert", self.A.nrow, self.size, nat.ptr(values),
                 nat.ptr(self._diagonal), nat.ptr(self.inverse_diagonal),
                 float(self.pivot_tolerance), nat.ptr(self._bad), nat.code(self.A.scalar_type),
                 nat.stream())

    def _check(self):
        if int(self._bad.item()):
            raise ValueError("Missing, nonfinite or singular/near-singular block pivot")

    def _apply(self, x, y, z, alpha, beta):
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        nat.call("wpc_block_diagonal_apply", self.A.nrow, self.size, nat.ptr(self.inverse_diagonal),
                 nat.ptr(x), nat.ptr(y), nat.ptr(z), float(alpha), float(beta),
                 nat.code(self.A.scalar_type), nat.stream())


class BlockILU0(BlockJacobi):
    """Synchronous parallel block ILU(0) with fixed Jacobi triangular sweeps
    (block_ilu.py:347-450).

    Each setup sweep recomputes every stored L/U block from the previous sweep on
    the original pattern (no fill); the apply runs ``solve_sweeps`` Jacobi steps
    for the unit-L and for the U solve from zero.  The result is a fixed linear,
    generally nonsymmetric preconditioner.
    """

    def __init__(self, A, factor_sweeps=12, solve_sweeps=4, relaxation=1.0, pivot_tolerance=None):
        if not isinstance(factor_sweeps, int) or factor_sweeps < 0:
            raise ValueError("factor_sweeps must be a nonnegative integer")
        if not isinstance(solve_sweeps, int) or solve_sweeps < 1:
            raise ValueError("solve_sweeps must be a positive integer")
        if not math.isfinite(relaxation) or not 0 < relaxation <= 1:
            raise ValueError("relaxation must be in (0,1]")
        super().__init__(A, pivot_tolerance)
        self.factor_sweeps, self.solve_sweeps = factor_sweeps, solve_sweeps
        self.relaxation = relaxation
        b = self.size
        vals = self.A.values.reshape(-1, b, b)
        self.factors = torch.empty_like(vals)
        other = torch.empty_like(vals)
        dt = nat.code(self.A.scalar_type)
        s = nat.stream()
        nnzb = self.A.nnz
        nat.call("wpc_ilu_initialize", nnzb, b, nat.ptr(self._rows), nat.ptr(self.A.columns),
                 nat.ptr(vals), nat.ptr(self.inverse_diagonal), nat.ptr(self.factors), dt, s)
        for _ in range(factor_sweeps):
            nat.call("wpc_ilu_sweep", nnzb, b, nat.ptr(self.A.offsets), nat.ptr(self.A.columns),
                     nat.ptr(self._rows), nat.ptr(vals), nat.ptr(self.factors),
                     nat.ptr(self.inverse_diagonal), nat.ptr(other), float(relaxation), dt, s)
            self.factors, other = other, self.factors
            self._invert(self.factors)
        nat.call("wpc_count_nonfinite", self.factors.numel(), nat.ptr(self.factors), dt,
                 nat.ptr(self._bad), s)
        self._check()
        self._buffers = [torch.empty(self.A.nrow * b, dtype=self.A.scalar_type, device=self.A.device)
                         for _ in range(4)]

    def _apply(self, x, y, z, alpha, beta):
        if alpha == 0.0:
            return BlockJacobi._apply(self, x, y, z, alpha, beta)
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        a, bb, c, d = self._buffers
        s = nat.stream()
        dt = nat.code(self.A.scalar_type)
        n = self.A.nrow
        off, col = nat.ptr(self.A.offsets), nat.ptr(self.A.columns)
        nat.call("wpc_fill", a.numel(), nat.ptr(a), dt, 0.0, s)
        for _ in range(self.solve_sweeps):
            nat.call("wpc_ilu_triangular", n, self.size, off, col, nat.ptr(self.factors),
                     nat.ptr(self.inverse_diagonal), nat.ptr(x), nat.ptr(a), nat.ptr(bb), 1, dt, s)
            a, bb = bb, a
        nat.call("wpc_fill", c.numel(), nat.ptr(c), dt, 0.0, s)
        for _ in range(self.solve_sweeps):
            nat.call("wpc_ilu_triangular", n, self.size, off, col, nat.ptr(self.factors),
                     nat.ptr(self.inverse_diagonal), nat.ptr(a), nat.ptr(c), nat.ptr(d), 0, dt, s)
            c, d = d, c
        nat.call("wpc_axpby", c.numel(), float(alpha), nat.ptr(c), float(beta), nat.ptr(y),
                 nat.ptr(z), dt, s)
