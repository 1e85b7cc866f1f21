"""Certified stopping for systems A = alpha*M + (positive semidefinite).

For such A the generalized eigenvalues of (A, M) are >= alpha, so the error
e = x - x* of any iterate obeys ||e||_M <= ||b - A x||_{M^-1} / alpha, and

    field_bound = ||b - A x||_{M^-1} / (alpha ||x||_M)

bounds the relative mass-weighted field error (exactly, up to the factor
1 / (1 - field_bound) from using ||x||_M for ||x*||_M).  ``TrueResidual``
recomputes b - A x with compensated float64 accumulation (wpc_dd_residual), so
the bound is not limited by the cancellation noise of a plain float64 product.
Both operator kinds of the reference's factored/assembled paths are covered:
``SparseOperator`` or a BSR matrix operator on an assembled float64 matrix
(r = b - A x) and the
identity-free ``SquaredLaplacianOperator`` (r = b - w M x - K diag(1/m) K x,
squared_laplacian.py:158-195).
"""

from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from .linear import _MatrixOperator
from .sparse import bsr_copy
from .sparse_operator import SparseOperator
from .squared_laplacian import SquaredLaplacianOperator


class TrueResidual:
    """Compensated r = b - A x and the norms ||r||^2_{M^-1}, ||x||^2_M on nvec
    interleaved float64 right-hand sides (all buffers preallocated)."""

    def __init__(self, op, mass: torch.Tensor | None, nvec: int):
        """mass: the M of the norms; the factored operator's own mass when None (its
        alpha*M term is the operator's, so any other mass is refused), ones (plain
        2-norms) for an assembled operator when None."""
        if isinstance(op, SquaredLaplacianOperator):
            if not op.identity or op.L.scalar_type != torch.float64:
                raise ValueError("TrueResidual needs an all-free float64 factored operator")
            if mass is None:
                mass = op.mass
            elif mass.data_ptr() != op.mass.data_ptr() and not torch.equal(mass, op.mass):
                raise ValueError("a factored operator's residual uses the operator's own mass")
            self.matrix, self.inverse, self.alpha = op.L, op.inverse_mass, op.mass_weight
            self.factored = True
        elif isinstance(op, SparseOperator):
            if op.scalar.scalar_type != torch.float64:
                raise ValueError("TrueResidual needs a float64 matrix")
            self.matrix, self.inverse, self.alpha = op.scalar, None, 0.0
            self.factored = False
        elif isinstance(op, _MatrixOperator):
            # a BSR matrix operator: the residual runs on its scalar CSR form
            if op.A.scalar_type != torch.float64:
                raise ValueError("TrueResidual needs a float64 matrix")
            self.matrix = op.A if op.A.block_shape == (1, 1) else bsr_copy(op.A, block_shape=(1, 1))
            self.inverse, self.alpha = None, 0.0
            self.factored = False
        else:
            raise TypeError("TrueResidual supports SparseOperator, BSR matrix operators and "
                            "SquaredLaplacianOperator")
        if not 1 <= nvec <= 8:
            raise ValueError("nvec must be in [1, 8]")
        n = self.matrix.nrow
        dev = self.matrix.device
        if mass is None:
            mass = torch.ones(n, dtype=torch.float64, device=dev)
        self.n, self.nvec, self.mass = n, nvec, mass
        size = n * nvec
        self.r = torch.empty(size, dtype=torch.float64, device=dev)
        self.t_hi = torch.empty(size, dtype=torch.float64, device=dev) if self.factored else None
        self.t_lo = torch.empty(size, dtype=torch.float64, device=dev) if self.factored else None
        self.partials = torch.empty(nat.lib().wpc_residual_partials(n) * 2 * nvec,
                                    dtype=torch.float64, device=dev)
        self.norms = torch.empty(2 * nvec, dtype=torch.float64, device=dev)

    def run(self, x: torch.Tensor, b: torch.Tensor, norms: torch.Tensor | None = None):
        """Launch on the current stream: self.r = b - A x (b may be self.r itself) and
        norms (default self.norms) = [||r||^2_{M^-1}, ||x||^2_M] per right-hand side."""
        A = self.matrix
        out = self.norms if norms is None else norms
        nat.call("wpc_dd_residual", self.n, nat.ptr(A.offsets), nat.ptr(A.columns),
                 nat.ptr(A.values), nat.ptr(self.mass), nat.ptr(self.inverse), float(self.alpha),
                 nat.ptr(x), nat.ptr(b), nat.ptr(self.t_hi), nat.ptr(self.t_lo), nat.ptr(self.r),
                 self.nvec, nat.ptr(self.partials), nat.ptr(out), nat.stream())
        return out


EPS = 2.0 ** -53


def witness_bounds(norms_next, norms_witness, norms_prev, nvec: int, alpha: float) -> np.ndarray:
    """Certificate for x' = fl(x + c), with c the correction computed from the
    residual r of x (the witness):

        ||x' - x*||_M <= ||r - A c||_{M^-1} / alpha + eps ||x'||_M

    since x + c - x* = c - A^-1 r and ||v||^2_M <= v^T A v / alpha <= ||A v||^2_{M^-1}
    / alpha^2 for A >= alpha M.  Unlike field_bounds(x'), this does not see the
    rounding noise of x' itself, which the lambda_min bound charges at alpha.
    norms_witness = [||r - A c||^2_{M^-1}, ...] (wpc_dd_residual with b = r, x = c),
    norms_prev = [||r||^2_{M^-1}, ...] of x (r stored rounded: slack eps ||r||),
    norms_next = [..., ||x'||^2_M]."""
    rho = np.sqrt(np.asarray(norms_witness, dtype=np.float64)[:nvec])
    r = np.sqrt(np.asarray(norms_prev, dtype=np.float64)[:nvec])
    x = np.sqrt(np.asarray(norms_next, dtype=np.float64)[nvec:2 * nvec])
    num = (1.0 + 4 * EPS) * rho + EPS * r
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(num == 0.0, 0.0, num / (alpha * x)) + 2 * EPS
    return np.where(np.isnan(out), np.inf, out)


def field_bounds(norms, nvec: int, alpha: float) -> np.ndarray:
    """Per-RHS bound ||r||_{M^-1} / (alpha ||x||_M) from host norms [2*nvec]."""
    norms = np.asarray(norms, dtype=np.float64)
    r = np.sqrt(norms[:nvec])
    x = np.sqrt(norms[nvec:2 * nvec])
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(r == 0.0, 0.0, r / (alpha * x))
    return np.where(np.isnan(out), np.inf, out)
