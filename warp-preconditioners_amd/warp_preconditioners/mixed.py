"""Mixed (indefinite) harmonic systems and their block preconditioners (drop-in
for mixed.py of the reference).

``MixedHarmonicSystem(K, mass, order, data_weight, equilibrate)`` assembles the
interleaved symmetric indefinite 2- or 3-field matrix (mixed.py:163-258):

* order 2: ``[[alpha M, -K], [-K, -M]]``, Schur complement ``alpha M + K M^-1 K``;
* order 3: ``[[alpha M, 0, -K], [0, -K, -M], [-K, -M, 0]]``, Schur complement
  ``alpha M - K M^-1 K M^-1 K`` (the benchmark's own signs, kept as they are);

plus the SPD companion ``H = K + alpha^(1/order) M``.  ``equilibrate=True``
builds ``D A D`` with ``D = (alpha^-1/4, alpha^1/4)`` or
``(alpha^-1/3, 1, alpha^1/3)``; ``rhs``/``solution``/``transform`` apply the change
of variables.  ``MatchingSchur`` eliminates the auxiliary fields with repeated
approximate inverses of H (mixed.py:292-322); ``ShiftedBlock`` is the first-order
block approximation with one H inverse per field (mixed.py:375-427).  Both are
linear operators for GMRES (the systems are indefinite: not CG).  The
per-vertex glue runs in the HIP library (``wpc_mixed_*``, ``wpc_shifted_*``);
products go through ``SparseOperator`` and the FSAI apply.
"""

from __future__ import annotations

import math

import torch

from . import _native as nat
from .fsai import FSAI
from .linear import LinearOperator
from .sparse import BlockType, BsrMatrix, bsr_copy, bsr_from_triplets
from .sparse_operator import SparseOperator


def _launch(name, *args):
    nat.call(name, *args, nat.stream())


class MixedHarmonicSystem:
    """The upstream benchmark's symmetric indefinite 2/3-field systems (mixed.py:163-290)."""

    def __init__(self, K, mass, order=2, data_weight=1.0, equilibrate=False):
        if not isinstance(K, BsrMatrix) or K.nrow != K.ncol or K.block_shape != (1, 1):
            raise ValueError("K must be square scalar BSR")
        if K.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("K must use float32 or float64")
        if (mass.dtype != K.scalar_type or mass.device != K.device
                or tuple(mass.shape) != (K.nrow,)):
            raise ValueError("mass must match K")
        if order not in (2, 3) or not math.isfinite(data_weight) or data_weight <= 0:
            raise ValueError("order must be 2 or 3 and data_weight finite positive")
        nat.require_device(K.values, "K")
        self.K = bsr_copy(K)
        self.mass = mass.clone()
        self.order, self.data_weight = order, data_weight
        self.vector_type = BlockType(1, order, K.scalar_type)
        self.input_data_weight = data_weight
        self.equilibrate = equilibrate
        self.scale = (1.0,) * order
        code = nat.code(K.scalar_type)
        n = K.nrow
        if equilibrate:
            root = data_weight ** (1 / order)
            self.scale = ((data_weight ** -0.25, data_weight ** 0.25) if order == 2
                          else (1 / root, 1.0, root))
            _launch("wpc_vector_scale", n, 1, nat.ptr(mass), nat.ptr(self.mass), root, 0.0, 0.0,
                    code)
            self.data_weight = 1.0
            data_weight = 1.0
        nnz = self.K.nnz
        count = n + nnz
        dev = K.device
        rows = torch.empty(count, dtype=torch.int32, device=dev)
        cols = torch.empty_like(rows)
        blocks = torch.empty((count, order, order), dtype=K.scalar_type, device=dev)
        hvalues = torch.empty(count, dtype=K.scalar_type, device=dev)
        bad = torch.zeros(1, dtype=torch.int32, device=dev)
        _launch("wpc_mixed_build", n, order, nat.ptr(self.K.offsets), nat.ptr(self.K.columns),
                nat.ptr(self.K.values), nat.ptr(self.mass), float(data_weight),
                float(data_weight ** (1 / order)), nat.ptr(rows), nat.ptr(cols), nat.ptr(blocks),
                nat.ptr(hvalues), nat.ptr(bad), code)
        if int(bad.item()):
            raise ValueError("mass must be finite positive")
        self.matrix = bsr_from_triplets(n, n, rows, cols, blocks)
        self.H = bsr_from_triplets(n, n, rows, cols, hvalues)

    def _check_vector(self, x, what="x must be an interleaved system vector"):
        n = self.K.nrow
        if (not isinstance(x, torch.Tensor) or x.dtype != self.mass.dtype
                or tuple(x.shape) != (n, self.order) or x.device != self.mass.device
                or not x.is_contiguous()):
            raise ValueError(what)

    def rhs(self, b):
        """Embed a scalar RHS as (b, 0[, 0]) per vertex (balanced when equilibrated)."""
        if b.dtype != self.mass.dtype or tuple(b.shape) != tuple(self.mass.shape) \
                or b.device != self.mass.device:
            raise ValueError("b must match mass")
        out = torch.zeros((self.K.nrow, self.order), dtype=b.dtype, device=b.device)
        out[:, 0] = b
        if self.equilibrate:
            self.transform(out, out)
        return out

    def transform(self, x, out=None):
        """Apply D: maps original RHS to balanced RHS, or balanced solution to original."""
        self._check_vector(x)
        if out is None:
            out = torch.empty_like(x)
        if out.dtype != x.dtype or out.shape != x.shape or out.device != x.device \
                or not out.is_contiguous():
            raise ValueError("out must match x")
        s = tuple(self.scale) + (0.0,) * (3 - self.order)
        _launch("wpc_vector_scale", x.numel(), self.order, nat.ptr(x), nat.ptr(out), *map(float, s),
                nat.code(x.dtype))
        return out

    def solution(self, x):
        """The primary field u of a (balanced) system solution."""
        self._check_vector(x)
        if self.equilibrate:
            x = self.transform(x)
        out = torch.empty_like(self.mass)
        _launch("wpc_mixed_component", self.K.nrow, self.order, nat.ptr(x), 0, None, nat.ptr(out),
                nat.code(x.dtype))
        return out


def _inverse_for(system, H_inverse, fsai_options):
    inverse = FSAI(system.H, **fsai_options) if H_inverse is None else H_inverse
    if (tuple(inverse.shape) != tuple(system.H.shape) or inverse.dtype != system.H.dtype
            or inverse.device != system.H.device):
        raise ValueError("H_inverse must match H")
    return inverse


class MatchingSchur(LinearOperator):
    """Block elimination with repeated approximate inverses of H = K + alpha^(1/p) M.

    The Schur inverse is approximated by ``H^-1 M H^-1`` (order 2) or
    ``-H^-1 M H^-1 M H^-1`` (order 3); the auxiliary fields follow by
    back-substitution through K and M^-1 (mixed.py:292-322).  No inner adaptive
    solve: a fixed linear operator, valid for right-preconditioned GMRES.  Owns
    scratch: do not apply one instance concurrently on independent streams.
    """

    def __init__(self, system, H_inverse=None, **fsai_options):
        self.system = system
        self.inverse = _inverse_for(system, H_inverse, fsai_options)
        self.K = SparseOperator(system.K, row_lanes=4)
        self._buffers = [torch.empty_like(system.mass) for _ in range(7)]
        super().__init__(system.matrix.shape, system.matrix.dtype, system.K.device, self._apply)

    def _apply(self, x, y, z, alpha, beta):
        s = self.system
        n, order, code = s.K.nrow, s.order, nat.code(s.mass.dtype)
        mass = nat.ptr(s.mass)
        t, v, w, g, u, a, lam = self._buffers
        px = nat.ptr(x)
        _launch("wpc_mixed_component", n, order, px, 1, mass, nat.ptr(t), code)
        self.K.matvec(t, v, v, 1.0, 0.0)
        if order == 3:
            _launch("wpc_mixed_component", n, order, px, 2, mass, nat.ptr(t), code)
            self.K.matvec(t, w, w, 1.0, 0.0)
            _launch("wpc_mass_multiply", n, nat.ptr(w), mass, 1, nat.ptr(t), code)
            self.K.matvec(t, w, w, 1.0, 0.0)
        _launch("wpc_mixed_rhs", n, order, px, nat.ptr(v), nat.ptr(w), nat.ptr(g), code)
        self.inverse.matvec(g, u, u, 1.0, 0.0)
        for step in range(order - 1):
            _launch("wpc_mass_multiply", n, nat.ptr(u), mass, 0, nat.ptr(t), code)
            self.inverse.matvec(t, u, u, -1.0 if order == 3 and step == 1 else 1.0, 0.0)
        self.K.matvec(u, v, v, 1.0, 0.0)
        _launch("wpc_mixed_auxiliary", n, order, px, 1 if order == 2 else 2, nat.ptr(v), mass,
                nat.ptr(a), code)
        if order == 3:
            self.K.matvec(a, v, v, 1.0, 0.0)
            _launch("wpc_mixed_auxiliary", n, order, px, 1, nat.ptr(v), mass, nat.ptr(lam), code)
        _launch("wpc_mixed_finish", n, order, nat.ptr(u), nat.ptr(a), nat.ptr(lam), nat.ptr(y),
                nat.ptr(z), float(alpha), float(beta), code)


class ShiftedBlock(LinearOperator):
    """First-order block approximation using H = K + alpha^(1/order) M (mixed.py:375-427).

    Order 2 uses the balanced block rotation (exact H inverses give preconditioned
    eigenvalues (1 +/- i t)/2, |t| <= 1); order 3 inverts the leading K blocks,
    with the resonance at generalized eigenvalue alpha^(1/3) left in place.
    Apply is capturable and allows input/output aliases; owns scratch.
    """

    def __init__(self, system, H_inverse=None, **fsai_options):
        self.system = system
        self.inverse = _inverse_for(system, H_inverse, fsai_options)
        self._buffers = [torch.empty_like(system.mass) for _ in range(6)]
        super().__init__(system.matrix.shape, system.matrix.dtype, system.K.device, self._apply)

    def _apply(self, x, y, z, alpha, beta):
        s = self.system
        n, order, code = s.K.nrow, s.order, nat.code(s.mass.dtype)
        a, b, c, u, v, w = self._buffers
        scale = float(s.data_weight ** 0.25)
        _launch("wpc_shifted_unpack", n, order, nat.ptr(x), nat.ptr(a), nat.ptr(b), nat.ptr(c),
                scale, code)
        self.inverse.matvec(a, u, u, 1.0, 0.0)
        self.inverse.matvec(b, v, v, 1.0, 0.0)
        if order == 3:
            self.inverse.matvec(c, w, w, 1.0, 0.0)
        _launch("wpc_shifted_finish", n, order, nat.ptr(u), nat.ptr(v), nat.ptr(w), nat.ptr(y),
                nat.ptr(z), scale, float(alpha), float(beta), code)
