"""First-order lifted biharmonic system (drop-in for biharmonic.py).

``BiharmonicSystem(K, mass, tau)`` builds the interleaved 2x2-block matrix
``[[M, -sqrt(tau) K], [sqrt(tau) K, M]]`` whose Schur complement is
``M + tau K M^-1 K`` and the SPD ``H = M + sqrt(tau) K`` for FSAI
(biharmonic.py:109-200).  ``RepeatedFSAI`` applies one scalar FSAI to both
components of interleaved vec2 arrays, i.e. ``(G (x) I2)^T (G (x) I2)``
(biharmonic.py:79-106), as a two-vector FSAI apply.  The lifted matrix is
nonsymmetric: use GMRES/BiCGSTAB, not CG/CR.
"""

from __future__ import annotations

import math

import torch

from . import _native as nat
from .fsai import FSAI
from .linear import LinearOperator
from .sparse import BlockType, BsrMatrix, bsr_copy, bsr_from_triplets


class RepeatedFSAI(LinearOperator):
    """Apply the same scalar FSAI independently to both components of vec2 arrays."""

    def __init__(self, factor: FSAI):
        self.factor = factor
        n = factor.G.nrow
        dt = factor.G.scalar_type if factor.row_scale2 is None else factor.source.scalar_type
        super().__init__((2 * n, 2 * n), BlockType(2, 2, dt), factor.device, self._apply)

    def _apply(self, x, y, z, alpha, beta):
        self.factor.apply_batch(x, y, z, alpha, beta, 2)


class BiharmonicSystem:
    """Lift a biharmonic solve to a real 2x2-block first-order BSR system."""

    def __init__(self, K, mass, tau=1.0):
        if not isinstance(K, BsrMatrix) or K.shape[0] != K.shape[1] or K.block_shape != (1, 1):
            raise ValueError("K must be a square scalar BSR matrix")
        if K.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("K must use float32 or float64")
        if (mass.dtype != K.scalar_type or mass.device != K.device
                or tuple(mass.shape) != (K.nrow,)):
            raise ValueError("mass must match K's scalar dtype, device, and row count")
        if not math.isfinite(tau) or tau <= 0:
            raise ValueError("tau must be finite and positive")
        nat.require_device(K.values, "K")
        K = bsr_copy(K)
        self.K, self.mass, self.tau = K, mass.clone(), tau
        self.vector_type = BlockType(1, 2, K.scalar_type)
        bad = torch.zeros(1, dtype=torch.int32, device=K.device)
        nat.call("wpc_count_nonpositive", K.nrow, nat.ptr(mass), nat.code(K.scalar_type),
                 nat.ptr(bad), nat.stream())
        if int(bad.item()):
            raise ValueError("mass must contain finite positive entries")
        n = K.nrow
        dev = K.device
        root = math.sqrt(tau)
        diag = torch.arange(n, dtype=torch.int64, device=dev)
        counts = (K.offsets[1:] - K.offsets[:-1]).to(torch.int64)
        krows = torch.repeat_interleave(diag, counts)
        kcols = K.columns.to(torch.int64)
        k = root * K.values
        rows = torch.cat([diag, krows])
        cols = torch.cat([diag, kcols])
        zeros_n = torch.zeros(n, dtype=K.scalar_type, device=dev)
        zeros_k = torch.zeros_like(k)
        blocks = torch.cat([
            torch.stack([torch.stack([self.mass, zeros_n], -1), torch.stack([zeros_n, self.mass], -1)], -2),
            torch.stack([torch.stack([zeros_k, -k], -1), torch.stack([k, zeros_k], -1)], -2),
        ])
        hvals = torch.cat([self.mass, k])
        self.matrix = bsr_from_triplets(n, n, rows, cols, blocks)
        self.H = bsr_from_triplets(n, n, rows, cols, hvals)

    def preconditioner(self, **fsai_options):
        return RepeatedFSAI(FSAI(self.H, **fsai_options))

    def rhs(self, b):
        """Embed a scalar RHS as interleaved (b, 0)."""
        if b.shape != self.mass.shape or b.dtype != self.mass.dtype or b.device != self.mass.device:
            raise ValueError("b must match mass's shape, dtype and device")
        out = torch.zeros((b.numel(), 2), dtype=b.dtype, device=b.device)
        out[:, 0] = b
        return out

    def solution(self, x, out=None):
        """Extract u from a lifted (n, 2) solution."""
        if (x.dtype != self.mass.dtype or tuple(x.shape) != (self.mass.numel(), 2)
                or x.device != self.mass.device):
            raise ValueError("x must be a vec2 solution on the system device")
        if out is None:
            out = torch.empty_like(self.mass)
        if out.dtype != self.mass.dtype or out.shape != self.mass.shape or out.device != self.mass.device:
            raise ValueError("out must match mass's shape, dtype and device")
        out.copy_(x[:, 0])
        return out
