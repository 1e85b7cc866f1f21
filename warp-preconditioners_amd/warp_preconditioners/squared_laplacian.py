"""Factored squared-Laplacian operator (drop-in for squared_laplacian.py).

``SquaredLaplacianOperator(L, mass, free_indices, row_lanes, mass_weight)`` applies
``(L diag(1/mass) L + mass_weight diag(mass))[free, free]`` as two sparse passes
and builds the Dirichlet right-hand side with the same arithmetic
(squared_laplacian.py:103-218).  All passes are ``wpc_sqlap_apply``.
"""

from __future__ import annotations

import math

import torch

from . import _native as nat
from .linear import LinearOperator
from .sparse import BsrMatrix, StreamPlan, bsr_copy

LANES = (1, 2, 4, 8, 16, 32)


class SquaredLaplacianOperator(LinearOperator):
    """Apply ``(L diag(1/mass) L + mass_weight diag(mass))[free,free]``.

    ``L`` is a symmetric scalar BSR matrix, ``mass`` positive lumped masses,
    ``free_indices`` distinct free scalar indices (the rest are constrained).
    The energy includes every row of L.  Instances own scratch; do not use one
    instance on concurrent streams.
    """

    def __init__(self, L, mass, free_indices, row_lanes=1, mass_weight=0.0):
        if not isinstance(L, BsrMatrix) or L.nrow != L.ncol or L.block_shape != (1, 1):
            raise ValueError("L must be a square scalar BSR matrix")
        if L.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("L must use float32 or float64")
        if mass.device != L.device or free_indices.device != L.device:
            raise ValueError("L, mass, and free_indices must be on the same device")
        if mass.dtype != L.scalar_type or mass.dim() != 1 or mass.numel() != L.nrow:
            raise ValueError("mass must match the scalar type and size of L")
        if free_indices.dtype != torch.int32 or free_indices.dim() != 1:
            raise ValueError("free_indices must be a one-dimensional int32 array")
        if not isinstance(row_lanes, int) or row_lanes not in LANES:
            raise ValueError("row_lanes must be one of 1, 2, 4, 8, 16, 32")
        if not math.isfinite(mass_weight) or mass_weight < 0:
            raise ValueError("mass_weight must be finite and nonnegative")
        nat.require_device(L.values, "L")
        self.mass_weight = float(mass_weight)
        self.L = bsr_copy(L)
        self.mass = mass
        self.free = free_indices.clone()
        n = L.nrow
        dev = L.device
        s = nat.stream()
        self.index = torch.empty(n, dtype=torch.int32, device=dev)
        nat.call("wpc_fill_i32", n, nat.ptr(self.index), -1, s)
        self.inverse_mass = torch.empty_like(mass)
        self.row_lanes = row_lanes
        status = torch.zeros(3, dtype=torch.int32, device=dev)
        nat.call("wpc_sqlap_index_free", self.free.numel(), nat.ptr(self.free), n,
                 nat.ptr(self.index), nat.ptr(status), s)
        nat.call("wpc_sqlap_mass_inverse", n, nat.ptr(mass), nat.ptr(self.inverse_mass),
                 nat.code(L.scalar_type), nat.ptr(status), s)
        errors = status.cpu()
        if errors[0] or errors[1]:
            raise ValueError("free_indices must be distinct and in bounds")
        if errors[2]:
            raise ValueError("mass entries must be finite and positive")
        # Every row free in natural order: the index maps are the identity.
        self.identity = self.free.numel() == n and bool(
            torch.equal(self.free, torch.arange(n, dtype=torch.int32, device=dev)))
        self._tmp = {}
        self._plan = None
        self._wm = None
        self._part_energy = None  # curvature partials of the fused CG pair (cg_update)
        nf = self.free.numel()
        super().__init__((nf, nf), L.dtype, dev, self._apply)

    def _scratch(self, dtype, count):
        t = self._tmp.get(dtype)
        if t is None or t.numel() < count:
            t = torch.empty(count, dtype=dtype, device=self.device)
            self._tmp[dtype] = t
            nat.bump_generation()
        return t

    @property
    def fused(self):
        return self.identity

    def _stream(self):
        if self._plan is None:
            self._plan = StreamPlan(self.L, None if self.row_lanes == 1 else self.row_lanes)
            self._wm = torch.empty_like(self.mass)
            nat.call("wpc_axpby", self.mass.numel(), self.mass_weight, nat.ptr(self.mass), 0.0,
                     None, nat.ptr(self._wm), nat.code(self.mass.dtype), nat.stream())
        return self._plan

    def product(self, x, y, z, alpha, beta, boundary=False, nvec=1, dot_w=None, dot_out=None):
        L = self.L
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        tmp = self._scratch(x.dtype, L.nrow * nvec)
        ident = self.identity and not boundary
        if ident and x.dtype in (torch.float32, torch.float64):
            # pass 1: tmp = M^-1 (L x); pass 2: z = alpha (L tmp + w M x) + beta y
            plan = self._stream()
            if alpha != 0.0:
                plan.run(x, None, tmp, 1.0, 0.0, nvec, s=self.inverse_mass)
            return plan.run(tmp, y, z, alpha, beta, nvec,
                            d=self._wm if self.mass_weight != 0.0 else None, u=x,
                            w=None if dot_w is None else dot_w.reshape(-1), out=dot_out)
        if dot_w is not None:
            raise ValueError("fused inner products need the identity free map")
        nat.call("wpc_sqlap_apply", L.nrow, nat.ptr(L.offsets), nat.ptr(L.columns),
                 nat.ptr(L.values), nat.code(L.scalar_type), nat.ptr(self.inverse_mass),
                 None if ident else nat.ptr(self.index), self.free.numel(),
                 None if ident else nat.ptr(self.free), nat.ptr(self.mass),
                 0.0 if boundary else self.mass_weight, nat.ptr(tmp), nat.ptr(x), nat.ptr(y),
                 nat.ptr(z), nvec, float(alpha), float(beta), int(boundary), self.row_lanes,
                 nat.stream())

    def cg_update_parts(self, nvec, dtype):
        """Partial count of `cg_update` for `nvec` vectors of `dtype`, or None when
        the fused pair does not apply (a free map other than the identity, vectors
        of another type than L, or a row mapping other than the sliced copy)."""
        if not self.identity or dtype != self.L.scalar_type:
            return None
        if dtype not in (torch.float32, torch.float64) or not 1 <= nvec <= 8:
            return None
        plan = self._stream()
        if plan.mode(nvec, dtype) != StreamPlan.SLICED:
            return None
        if not plan.sliced_fits(nvec, torch.finfo(dtype).bits // 8):
            return None
        return nat.lib().wpc_sqlap_cg_partials(self.L.nrow)

    def energy_partials(self, nvec):
        """The buffer of the first fused pass's curvature partials
        (`cg_update_parts` per vector), sized for 8 vectors when first made, so a
        graph captured for fewer vectors never sees it move."""
        count = nat.lib().wpc_sqlap_cg_partials(self.L.nrow)
        if self._part_energy is None or self._part_energy.numel() < count * nvec:
            self._part_energy = torch.zeros(max(count * max(nvec, 8), 1), dtype=torch.float64,
                                            device=self.L.device)
            nat.bump_generation()
        return self._part_energy

    def cg_update(self, p, r, nvec, state, part_rr, passes=3, x=None, t=None):
        """The operator's two passes on p inside fused CG with the residual update
        folded into the second (`wpc_sqlap_cg_update`): r -= ALPHA A p with ALPHA =
        RZ / p.Ap from `state`, A p never stored; leaves `cg_update_parts` r.r
        partials in `part_rr` and the step's scalars in `state`; with `x` the
        second pass also advances x += ALPHA p.  `passes` 1 or 2 runs one pass alone
        (timing; 2 also after `direction_energy`); 6 runs the second pass with ALPHA
        taken from `state` (as wpc_cg_alpha leaves it after folding
        `energy_partials`), x required.  `t` holds M^-1 L p between the passes (the
        operator's scratch when None)."""
        plan = self._stream()
        sp, sl, sc, sv = plan._sliced()
        count = nat.lib().wpc_sqlap_cg_partials(self.L.nrow)
        part_e = self.energy_partials(nvec)
        tmp = self._scratch(p.dtype, self.L.nrow * nvec) if t is None else t
        nat.call("wpc_sqlap_cg_update", self.L.nrow, nat.ptr(sp), nat.ptr(sl),
                 nat.ptr(plan._sell[7]), nat.ptr(plan._sell[5]), nat.ptr(sc), int(plan._sell[6]),
                 nat.ptr(sv), nat.code(sv.dtype), nat.ptr(self.inverse_mass),
                 nat.ptr(self._wm) if self.mass_weight != 0.0 else None, nat.ptr(p), nat.ptr(tmp),
                 nat.ptr(r), nat.code(p.dtype), nvec, nat.ptr(part_e), nat.ptr(part_rr),
                 nat.ptr(state), int(passes), nat.ptr(x), nat.stream())
        return count

    def direction_energy(self, z, p, t, nvec, nparts_rz, part_rz, part_rr, state):
        """The previous CG step's direction step folded into this step's first pass
        (`wpc_sqlap_cg_direction_energy`): BETA from the r.z and r.r partials,
        p = z + BETA p, t = M^-1 L z + BETA t (t as the previous first pass left
        it), curvature partials for `cg_update(..., passes=2, t=t)`.  With
        `part_rz` None, BETA as wpc_cg_beta left it in `state` (nothing folded; then
        wpc_cg_alpha and `cg_update(..., passes=6, t=t)` follow)."""
        plan = self._stream()
        sp, sl, sc, sv = plan._sliced()
        part_e = self.energy_partials(nvec)
        nat.call("wpc_sqlap_cg_direction_energy", self.L.nrow, nat.ptr(sp), nat.ptr(sl),
                 nat.ptr(plan._sell[7]), nat.ptr(plan._sell[5]), nat.ptr(sc), int(plan._sell[6]),
                 nat.ptr(sv), nat.code(sv.dtype), nat.ptr(self.inverse_mass),
                 nat.ptr(self._wm) if self.mass_weight != 0.0 else None, nat.ptr(z), nat.ptr(p),
                 nat.ptr(t), nat.code(p.dtype), nvec, int(nparts_rz), nat.ptr(part_rz),
                 nat.ptr(part_rr), nat.ptr(part_e), nat.ptr(state), nat.stream())

    def _apply(self, x, y, z, alpha, beta):
        self.product(x, y, z, alpha, beta, False)

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        return self.product(x, y, z, alpha, beta, False, nvec, dot_w, dot_out)

    def rhs(self, boundary_values, out=None):
        if (boundary_values.dtype != self.scalar_type or boundary_values.device != self.device
                or boundary_values.dim() != 1 or boundary_values.numel() != self.L.nrow):
            raise ValueError("boundary_values must match the scalar type, device, and size of L")
        nf = self.free.numel()
        if out is None:
            out = torch.empty(nf, dtype=self.scalar_type, device=self.device)
        elif (out.dtype != self.scalar_type or out.device != self.device or out.dim() != 1
              or out.numel() != nf):
            raise ValueError("out must match the reduced operator")
        self.product(boundary_values, out, out, -1.0, 0.0, True)
        return out
