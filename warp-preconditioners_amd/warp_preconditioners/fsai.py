"""Adaptive, diagonally equilibrated FSAI on gfx950 (drop-in for fsai.py).

``FSAI(A, ...)`` builds a lower-triangular sparse factor G with positive diagonal
so that ``G^T G`` approximates ``A^-1``, and applies ``z = alpha*G^T G x + beta*y``
as two sparse passes through an explicitly stored transpose.  Construction,
validation messages and the factor pipeline follow the reference's FSAI.__init__
(fsai.py:222-344): validate -> scalarize -> diagonal -> build -> status read-back
-> pack -> optional conversion (+validation) -> transpose -> optional refit plan.
Every device step is a HIP kernel from include/wpc.h.

Beyond the reference, ``factor_dtype`` may be torch.float16 or torch.bfloat16 for a
float32/float64 matrix: the factor is then stored row-scaled, q = G / s with
s_i = max_j |G_ij| and s_i^2 kept in float32 (``row_scale2``), and applied as
q^T diag(s^2) q with accumulation in the vector type.
"""

from __future__ import annotations

import torch

from . import _native as nat
from .linear import LinearOperator
from .sparse import BsrMatrix, StreamPlan, bsr_transposed, scalarize

LANES = (1, 2, 4, 8, 16, 32)
_FLOATS = (torch.float16, torch.float32, torch.float64)
DEFAULT_PIVOT_FLOOR = {torch.float16: 1.0e-3, torch.float32: 1.0e-6, torch.float64: 1.0e-12}


def _pack_factor(n: int, width: int, cols: torch.Tensor, vals: torch.Tensor) -> BsrMatrix:
    """fsai_pack.py:77-90: sort padded rows, scan counts, compact into CSR."""
    dev = vals.device
    counts = torch.empty(n + 1, dtype=torch.int32, device=dev)
    s = nat.stream()
    nat.call("wpc_fsai_sort_rows", n, width, nat.ptr(cols), nat.ptr(vals), nat.code(vals.dtype),
             nat.ptr(counts), s)
    offsets = torch.empty(n + 1, dtype=torch.int32, device=dev)
    nat.call("wpc_scan_inclusive", n + 1, nat.ptr(counts), nat.ptr(offsets), s)
    nnz = int(offsets[-1].item()) if n > 0 else 0
    out_cols = torch.empty(nnz, dtype=torch.int32, device=dev)
    out_vals = torch.empty(nnz, dtype=vals.dtype, device=dev)
    nat.call("wpc_fsai_pack", n, width, nat.ptr(offsets), nat.ptr(cols), nat.ptr(vals),
             nat.code(vals.dtype), nat.ptr(out_cols), nat.ptr(out_vals), s)
    return BsrMatrix(n, n, (1, 1), offsets, out_cols, out_vals)


def build_factor(S: BsrMatrix, width: int, step: int, tolerance: float, pivot_floor: float):
    """Diagonal + adaptive build on a scalar matrix (fsai.py:284-312).

    Returns (scale, padded columns, padded values, host status[3])."""
    n = S.nrow
    dev = S.device
    dt = S.scalar_type
    scale = torch.empty(n, dtype=dt, device=dev)
    status = torch.zeros(3, dtype=torch.int32, device=dev)
    s = nat.stream()
    nat.call("wpc_fsai_diagonal", n, nat.ptr(S.offsets), nat.ptr(S.columns), nat.ptr(S.values),
             nat.code(dt), nat.ptr(scale), nat.ptr(status), s)
    cols = torch.empty(n * width, dtype=torch.int32, device=dev)
    vals = torch.empty(n * width, dtype=dt, device=dev)
    nat.call("wpc_fsai_build", n, nat.ptr(S.offsets), nat.ptr(S.columns), nat.ptr(S.values),
             nat.code(dt), nat.ptr(scale), float(tolerance), float(pivot_floor), width, step,
             nat.ptr(cols), nat.ptr(vals), nat.ptr(status), s)
    return scale, cols, vals, status.cpu()


class FSAI(LinearOperator):
    """Approximate inverse ``G.T @ G`` of an SPD BSR matrix (reference fsai.py:199-381).

    ``max_row_size`` includes the diagonal; each search adds up to ``max_step_size``
    distinct largest-residual entries of the lower-triangular graph frontier;
    ``kap_tolerance`` stops a row when the relative energy gain drops below it.
    Setup synchronizes (status read-back); apply is stream-ordered and capturable.
    Instances own scratch: do not apply one instance on concurrent streams.
    """

    def __init__(self, A, max_row_size=8, kap_tolerance=1.0e-3, pivot_floor=None,
                 apply_lanes=1, factor_dtype=None, reuse_pattern=False, max_step_size=1):
        if (not isinstance(A, BsrMatrix) or A.shape[0] != A.shape[1]
                or A.block_shape[0] != A.block_shape[1]):
            raise ValueError("FSAI requires a square BSR matrix with square blocks")
        if A.scalar_type not in _FLOATS:
            raise TypeError("FSAI supports float16, float32 and float64")
        if not isinstance(max_row_size, int) or not 1 <= max_row_size <= 64:
            raise ValueError("max_row_size must be an integer in [1, 64]")
        if not isinstance(max_step_size, int) or not 1 <= max_step_size <= 64:
            raise ValueError("max_step_size must be an integer in [1, 64]")
        if not 0 <= kap_tolerance < 1:
            raise ValueError("kap_tolerance must be in [0, 1)")
        if pivot_floor is None:
            pivot_floor = DEFAULT_PIVOT_FLOOR[A.scalar_type]
        if not 0 < pivot_floor < 1:
            raise ValueError("pivot_floor must be in (0, 1)")
        if not isinstance(apply_lanes, int) or apply_lanes not in LANES:
            raise ValueError("apply_lanes must be one of 1, 2, 4, 8, 16, 32")
        if factor_dtype is None:
            factor_dtype = A.scalar_type
        scaled = (factor_dtype in (torch.float16, torch.bfloat16)
                  and A.scalar_type in (torch.float32, torch.float64))
        if factor_dtype != A.scalar_type and not scaled and not (
            A.scalar_type == torch.float64 and factor_dtype == torch.float32
        ):
            raise ValueError(
                "factor_dtype must be the matrix scalar type, or wp.float32 for a float64 matrix"
            )
        if reuse_pattern and A.row_counts is not None:
            raise ValueError(
                "reuse_pattern requires compact BSR storage; canonicalize with bsr_copy"
            )
        nat.require_device(A.values, "A")
        self.max_row_size = max_row_size
        self.max_step_size = max_step_size
        self.factor_dtype = factor_dtype
        self.apply_lanes = apply_lanes
        self.kap_tolerance = kap_tolerance
        self.pivot_floor = pivot_floor
        self.source = A
        A.nnz_sync()
        if A.block_shape == (1, 1) and not reuse_pattern:
            scalar = A
        else:
            scalar = scalarize(A)  # keeps stored zeros, as scalarize_compact
        n = scalar.nrow
        dtype, device = scalar.scalar_type, scalar.device
        _, cols, vals, diagnostics = build_factor(scalar, max_row_size, max_step_size,
                                                  kap_tolerance, pivot_floor)
        if diagnostics[0]:
            raise ValueError(
                f"FSAI requires finite positive diagonals ({int(diagnostics[0])} invalid rows)"
            )
        self.truncated_rows = int(diagnostics[1])
        if diagnostics[2]:
            raise ValueError("FSAI factor lost finite entries or a positive diagonal")
        self.G = _pack_factor(n, max_row_size, cols, vals)
        self.row_scale2 = None
        if factor_dtype != dtype:
            status = torch.zeros(1, dtype=torch.int32, device=device)
            values = torch.empty(self.G.nnz, dtype=factor_dtype, device=device)
            if scaled:
                self.row_scale2 = torch.empty(n, dtype=torch.float32, device=device)
                nat.call("wpc_fsai_compress_scaled", n, nat.ptr(self.G.offsets),
                         nat.ptr(self.G.columns), nat.ptr(self.G.values), nat.code(dtype),
                         nat.ptr(values), nat.code(factor_dtype), nat.ptr(self.row_scale2),
                         nat.ptr(status), nat.stream())
            else:
                nat.call("wpc_convert", self.G.nnz, nat.ptr(self.G.values), nat.code(dtype),
                         nat.ptr(values), nat.code(factor_dtype), nat.stream())
                nat.call("wpc_fsai_validate_factor", n, nat.ptr(self.G.offsets),
                         nat.ptr(self.G.columns), nat.ptr(values), nat.code(factor_dtype),
                         nat.ptr(status), nat.stream())
            self.G = BsrMatrix(n, n, (1, 1), self.G.offsets, self.G.columns, values)
            if int(status.item()):
                raise ValueError(
                    "factor_dtype conversion lost finite entries or a positive diagonal"
                )
        self.GT = bsr_transposed(self.G)
        self._plans = None
        self._tmp = {}
        super().__init__(A.shape, A.dtype, device, self._apply)
        self._scalar_n = n
        self._refit = None
        if reuse_pattern:
            from .fsai_reuse import RefitPlan

            self._refit = RefitPlan(self, A, scalar, max_row_size, pivot_floor)

    # ------------------------------------------------------------------ refit
    def update(self, A=None):
        """Refit on the original supports, preserving the G/GT buffers (fsai.py:346-362)."""
        if self._refit is None:
            raise ValueError("Construct FSAI with reuse_pattern=True before updating")
        A = self.source if A is None else A
        if not isinstance(A, BsrMatrix):
            raise ValueError("update requires a Warp BSR matrix (pass A after release_source)")
        self._refit.update(self, A)
        self.source = A
        return self

    def release_source(self):
        """Drop the reference to the source matrix (frees it when nothing else holds
        it); update() and quality() then need an explicit A."""
        self.source = None
        return self

    def quality(self, A=None, probes=4, seed=17):
        """Rademacher estimate of ||G A G^T - I||_F / sqrt(n) (fsai.py:364-372)."""
        from .fsai_reuse import quality

        return quality(self, self.source if A is None else A, probes, seed)

    def dense_factor(self):
        """Host copy of the effective factor (row scales applied), for inspection."""
        g = self.G.to_dense()
        if self.row_scale2 is not None:
            g = g * self.row_scale2.double().sqrt().cpu().numpy()[:, None]
        return g

    # ------------------------------------------------------------------ apply
    def _scratch(self, dtype, count):
        t = self._tmp.get(dtype)
        if t is None or t.numel() < count:
            t = torch.empty(count, dtype=dtype, device=self.device)
            self._tmp[dtype] = t
            nat.bump_generation()
        return t

    # float32/float64 vectors go through wpc_csr_stream: apply_lanes == 1 (the
    # default) lets the plan's autotuner pick the row mapping, other values force
    # that many lanes per row; binary16 vectors use the multi-lane kernel.
    fused = True

    def plans(self):
        if self._plans is None:
            lanes = None if self.apply_lanes == 1 else self.apply_lanes
            self._plans = (StreamPlan(self.G, lanes), StreamPlan(self.GT, lanes))
        return self._plans

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        """z = alpha G^T G x + beta y on nvec interleaved right-hand sides; with
        ``dot_w`` also dot_out[v] = sum_i dot_w[i,v] z[i,v] (device, float64)."""
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        n = self._scalar_n
        tmp = self._scratch(x.dtype, n * nvec)
        streamed = x.dtype in (torch.float32, torch.float64)
        if streamed:
            pg, pgt = self.plans()
            if alpha != 0.0:
                pg.run(x, None, tmp, 1.0, 0.0, nvec, s=self.row_scale2)
            return pgt.run(tmp, y, z, alpha, beta, nvec,
                           w=None if dot_w is None else dot_w.reshape(-1), out=dot_out)
        nat.call("wpc_fsai_apply", n, nat.ptr(self.G.offsets), nat.ptr(self.G.columns),
                 nat.ptr(self.G.values), nat.ptr(self.GT.offsets), nat.ptr(self.GT.columns),
                 nat.ptr(self.GT.values), nat.code(self.G.scalar_type), nat.ptr(self.row_scale2),
                 nat.ptr(tmp), nat.ptr(x), nat.ptr(y), nat.ptr(z), nat.code(x.dtype), nvec,
                 float(alpha), float(beta), self.apply_lanes, nat.stream())

    def _apply(self, x, y, z, alpha, beta):
        # vec_b arrays are viewed as interleaved scalars (fsai.py:375-377)
        self.apply_batch(x, y, z, alpha, beta, 1)
