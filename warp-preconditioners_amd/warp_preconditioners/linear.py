"""Linear-operator protocol and device-resident CG / CR drivers.

The reference passes its operators to ``warp.optim.linear`` (LinearOperator,
aslinearoperator, cg, cr, preconditioner; e.g. tests/test_fsai.py:64,
benchmarks/synthetic.py:50, fsai_reuse.py:256).  Warp has no ROCm backend, so
this module provides the same call signatures on torch device buffers, with
every vector step in the HIP library (include/wpc.h, Krylov section):

* ``matvec(x, y, z, alpha, beta)`` computes ``z = alpha * Op x + beta * y``.
* ``cg``/``cr`` stop when ``||r||_2 <= max(atol, tol * ||b||_2)``, checked on
  the host every ``check_every`` iterations; a step whose incoming residual
  already meets the target changes nothing, so the returned count is the first
  converged iteration rounded up to a multiple of ``check_every`` (with
  ``check_every=0`` the exact count is returned as a device tensor, like Warp's
  capturable loop).  Chunks of ``check_every`` iterations are replayed from a
  captured HIP graph when ``use_cuda_graph``.
"""

from __future__ import annotations

import math

import numpy as np

import torch

from . import _native as nat
from .sparse import BlockType, BsrMatrix, StreamPlan, _product, time_graph


class LinearOperator:
    def __init__(self, shape, dtype, device, matvec):
        self.shape = tuple(shape)
        self.dtype = dtype
        self.device = torch.device(device) if device is not None else None
        self.matvec = matvec

    @property
    def scalar_type(self):
        return self.dtype.scalar_type if isinstance(self.dtype, BlockType) else self.dtype


class _MatrixOperator(LinearOperator):
    """BsrMatrix as an operator: Warp's bsr_mv with aliasing handled.

    Scalar CSR runs through a StreamPlan.  3x3 blocks (float32/float64 values) run
    the sliced block product (BsrPlan) or the scalarised CSR copy through a
    StreamPlan, whichever is faster for the (vector count, vector dtype): the
    choice is timed once outside graph capture (the block product is kept unless
    the scalar one is more than BLOCK_MARGIN faster) and shared by every operator
    over the same index buffers, as the reference's elasticity benchmark picks
    between bsr_mv and the cooperative CSR operator (benchmarks/elasticity.py:
    155-171).  Other blocks use the generic block kernel."""

    BLOCK_MARGIN = 0.03
    _block_chosen = {}

    def __init__(self, A: BsrMatrix, lanes=1):
        self.A = A
        self.lanes = lanes
        self._alias = None
        self._plan = None
        self._bsr = None
        self._scalar_plan = None
        self.block_timings = {}
        super().__init__(A.shape, A.dtype, A.device, self._apply)

    @property
    def block3(self):
        return self.A.block_shape == (3, 3) and self.A.scalar_type in (torch.float32, torch.float64)

    @property
    def fused(self):
        return self.A.block_shape == (1, 1) or self.block3

    def _apply(self, x, y, z, alpha, beta):
        xf, yf, zf = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        if xf.data_ptr() == zf.data_ptr():
            if self._alias is None or self._alias.numel() != xf.numel() or self._alias.dtype != xf.dtype:
                self._alias = torch.empty_like(xf)
            self._alias.copy_(xf)
            xf = self._alias
        if xf.dtype in (torch.float32, torch.float64) and (
                self.A.block_shape == (1, 1) or (self.block3 and self._block_vectors(xf, 1))):
            self.apply_batch(xf, yf, zf, alpha, beta, 1)
        else:
            _product(self.A, xf, yf, zf, alpha, beta, 8)

    # ---------------------------------------------------------------- 3x3 blocks
    def _block_vectors(self, x, nvec):
        """Vectors the block paths take: float64 (any value type) or float32 with
        float32 values, up to 8 (4 for the block product) interleaved."""
        return nvec <= 8 and (x.dtype == torch.float64 or
                              (x.dtype == torch.float32 and self.A.scalar_type == torch.float32))

    def _bsr_plan(self):
        if self._bsr is None:
            from .sparse import BsrPlan

            self._bsr = BsrPlan(self.A)
        return self._bsr

    def _scalar(self):
        if self._scalar_plan is None:
            from .sparse import bsr_copy

            self._scalar_plan = StreamPlan(bsr_copy(self.A, block_shape=(1, 1)),
                                           None if self.lanes == 1 else self.lanes)
        return self._scalar_plan

    def block_mode(self, nvec, dtype):
        """'bsr' (sliced block product) or 'csr' (scalarised copy) for this count."""
        if nvec > 4:
            return "csr"
        A = self.A
        key = (A.offsets.data_ptr(), A.columns.data_ptr(), A.nrow, A.nnz, A.scalar_type, nvec, dtype)
        mode = _MatrixOperator._block_chosen.get(key)
        if mode is not None:
            return mode
        if torch.cuda.is_current_stream_capturing():
            return "bsr"
        n = A.shape[0]
        x = torch.ones(n * nvec, dtype=dtype, device=A.device)
        z = torch.empty_like(x)
        bsr = self._bsr_plan()
        if not bsr.fits(x, nvec):
            mode = "csr"
        else:
            scalar = self._scalar()
            t_bsr = time_graph(lambda: bsr.run(x, None, z, 1.0, 0.0, nvec))
            t_csr = time_graph(lambda: scalar.run(x, None, z, 1.0, 0.0, nvec))
            self.block_timings[(nvec, dtype)] = dict(bsr=t_bsr, csr=t_csr)
            mode = "bsr" if t_bsr <= t_csr * (1 + self.BLOCK_MARGIN) else "csr"
            if mode == "bsr":
                self._scalar_plan = None  # the scalar copy is no longer needed
        _MatrixOperator._block_chosen[key] = mode
        return mode

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        xf, yf, zf = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        if xf.data_ptr() == zf.data_ptr():
            xf = xf.clone()
        A = self.A
        w = None if dot_w is None else dot_w.reshape(-1)
        if self.block3:
            if not self._block_vectors(xf, nvec):
                raise NotImplementedError("block products need float64 vectors (or float32 "
                                          "vectors with float32 values), at most 8")
            if self.block_mode(nvec, xf.dtype) == "bsr":
                return self._bsr_plan().run(xf, yf, zf, alpha, beta, nvec, w=w, out=dot_out)
            return self._scalar().run(xf, yf, zf, alpha, beta, nvec, w=w, out=dot_out)
        if A.block_shape != (1, 1):
            raise NotImplementedError("batched products need a scalar or 3x3 block matrix")
        if xf.dtype in (torch.float32, torch.float64):
            if self._plan is None:
                self._plan = StreamPlan(A, None if self.lanes == 1 else self.lanes)
            return self._plan.run(xf, yf, zf, alpha, beta, nvec, w=w, out=dot_out)
        nat.call("wpc_csr_spmv", A.nrow, nat.ptr(A.offsets), nat.ptr(A.columns), nat.ptr(A.values),
                 nat.code(A.scalar_type), None, nat.ptr(xf), nat.ptr(yf), nat.ptr(zf),
                 nat.code(xf.dtype), nvec, float(alpha), float(beta), self.lanes, nat.stream())


def aslinearoperator(A) -> LinearOperator:
    if A is None or isinstance(A, LinearOperator):
        return A
    if isinstance(A, BsrMatrix):
        return _MatrixOperator(A)
    raise TypeError("expected a LinearOperator or BsrMatrix")


def preconditioner(A: BsrMatrix, ptype: str = "diag") -> LinearOperator:
    """Scalar Jacobi (``ptype='diag'``) or identity (``'id'``)."""
    if ptype == "id":
        return None
    if ptype != "diag":
        raise ValueError(f"unsupported preconditioner type {ptype!r}")
    from .block_ilu import BlockJacobi
    from .sparse import bsr_copy

    scalar = A if A.block_shape == (1, 1) else bsr_copy(A, block_shape=(1, 1))
    jac = BlockJacobi(scalar)
    op = LinearOperator(A.shape, A.dtype, A.device, lambda x, y, z, a, b: jac.matvec(
        x.reshape(-1), y.reshape(-1), z.reshape(-1), a, b))
    op.inner = jac
    return op


# Largest nparts * nvec whose scalar fold is repeated by every workgroup of the
# following vector kernel instead of running as a kernel of its own.
FOLD_IN_PLACE = 4096
# Fold CG's update step into the operator's last pass where the operator offers it
# (SquaredLaplacianOperator.cg_update); False keeps the separate update launch.  Read
# when a KrylovWork is made (its `fuse_update`).
FUSE_OPERATOR_UPDATE = True
# Inside a run of fused CG steps (KrylovWork.steps), fold each step's direction
# step into the next step's first operator pass (recurrence form: four launches
# per step instead of five); the run's last step still ends with the direction
# step, so the state between runs is the same.  Read when a KrylovWork is made.
CG_RECURRENCE = True


class KrylovWork:
    """Device buffers and launch sequence of one CG/CR solve on nvec interleaved RHS."""

    def __init__(self, method, A, M, n, nvec, dtype, device):
        self.method = 0 if method == "cg" else 1
        self.A, self.M = A, M
        self.n, self.nvec, self.dtype = n, nvec, dtype
        size = n * nvec
        new = lambda: torch.empty(size, dtype=dtype, device=device)  # noqa: E731
        self.r, self.p, self.ap = new(), new(), new()
        self.z = new() if (M is not None or self.method == 1) else self.r
        self.map = new() if self.method == 1 else None
        self.az = new() if self.method == 1 else None
        self.P = nat.lib().wpc_krylov_partials(n)
        self.parts = torch.zeros(3, self.P * nvec, dtype=torch.float64, device=device)
        self.state = torch.zeros(nat_state_fields() * nvec, dtype=torch.float64, device=device)
        self.dot_parts = torch.empty(nat.lib().wpc_dot_partials(n) * nvec, dtype=torch.float64,
                                     device=device)
        self.code = nat.code(dtype)
        self.graph = None
        self.graph_steps = 0
        self._graph_key = None
        self._m_parts = None  # the preconditioner's partial count (fused steps)
        self._rr_parts = None  # r.r partials of an operator pass with the update folded in
        # Fold the update into the operator's last pass where the operator offers it.
        # Off for works that also advance in grouped launches (grouped.py), whose
        # batched kernels keep the separate update: both then take the same steps.
        self.fuse_update = FUSE_OPERATOR_UPDATE
        self.recurrence = CG_RECURRENCE
        self._m_part = None  # the preconditioner's r.z partials of its last pass
        self.use_fused = (self.method == 0 and M is not None and bool(getattr(A, "fused", False))
                          and bool(getattr(M, "fused", False)) and hasattr(A, "apply_batch")
                          and hasattr(M, "apply_batch"))

    # -- helpers
    def _field(self, k):
        return self.state[k * self.nvec:(k + 1) * self.nvec]

    def _dot(self, a, b, out):
        nat.call("wpc_dot", self.n, self.nvec, nat.ptr(a), nat.ptr(b), self.code,
                 nat.ptr(self.dot_parts), nat.ptr(out), nat.stream())

    def _op(self, op, x, y, z, alpha, beta):
        if self.nvec == 1:
            op.matvec(x, y, z, alpha, beta)
        else:
            op.apply_batch(x, y, z, alpha, beta, self.nvec)

    def _apply_M(self, src, dst):
        if self.M is None:
            if dst.data_ptr() != src.data_ptr():
                dst.copy_(src)
        else:
            self._op(self.M, src, dst, dst, 1.0, 0.0)

    def start(self, b, x, tol, atol):
        """r = b - A x; initial direction; device-side target."""
        self._op(self.A, x, b, self.r, -1.0, 1.0)
        self._dot(b, b, self._field(2))  # ||b||^2 into TARGET
        nat.call("wpc_krylov_set_target", self.nvec, nat.ptr(self.state), float(tol), float(atol),
                 nat.stream())
        self._dot(self.r, self.r, self._field(1))
        self._apply_M(self.r, self.z)
        if self.method == 0:
            self._dot(self.r, self.z, self._field(0))
            self.p.copy_(self.z)
        else:
            self._op(self.A, self.z, self.az, self.az, 1.0, 0.0)
            self._dot(self.z, self.az, self._field(0))
            self.p.copy_(self.z)
            self.ap.copy_(self.az)

    def step(self, x):
        s = nat.stream()
        pd, pr, pz = (nat.ptr(self.parts[k]) for k in range(3))
        st = nat.ptr(self.state)
        if self.use_fused and self._fused_update_step(x, s, st):
            return
        if self.use_fused:
            # A p (+p.Ap partials) | alpha | update (+r.r partials) | M r (+r.z partials)
            # | beta | direction
            part, nparts = self.A.apply_batch(self.p, self.ap, self.ap, 1.0, 0.0, self.nvec,
                                              dot_w=self.p)
            fold_a = nparts * self.nvec <= FOLD_IN_PLACE
            # With both steps folded, x advances in the direction step (x += ALPHA p
            # before p is replaced) and the update reads only r and Ap.  Whether the
            # direction step folds depends on the preconditioner's partial count, known
            # from its previous application (the first step advances x in the update;
            # a folded direction step is exact for any count, only slower for large).
            xd = fold_a and self._m_parts is not None and \
                max(self._m_parts, self.P) * self.nvec <= FOLD_IN_PLACE
            if fold_a:
                nat.call("wpc_cg_update_fold", self.n, self.nvec, None if xd else nat.ptr(x),
                         nat.ptr(self.r), nat.ptr(self.p), nat.ptr(self.ap), self.code, nparts,
                         nat.ptr(part), st, pr, s)
            else:
                nat.call("wpc_cg_alpha", self.nvec, nparts, nat.ptr(part), st, s)
                nat.call("wpc_cg_update", self.n, self.nvec, nat.ptr(x), nat.ptr(self.r),
                         nat.ptr(self.p), nat.ptr(self.ap), self.code, st, pr, s)
            part, nparts = self.M.apply_batch(self.r, self.z, self.z, 1.0, 0.0, self.nvec,
                                              dot_w=self.r)
            self._m_parts, self._m_part = nparts, part
            if xd or (fold_a and max(nparts, self.P) * self.nvec <= FOLD_IN_PLACE):
                nat.call("wpc_cg_direction_fold", self.n, self.nvec, nat.ptr(self.z),
                         nat.ptr(self.p), nat.ptr(x) if xd else None, self.code, nparts,
                         nat.ptr(part), self.P, pr, st, s)
            else:
                # (after a fused update too: it leaves RZ as the scalar fold expects)
                nat.call("wpc_cg_beta", self.nvec, nparts, nat.ptr(part), self.P, pr, st, s)
                nat.call("wpc_cg_direction", self.n, self.nvec, nat.ptr(self.z), nat.ptr(self.p),
                         self.code, st, s)
            return
        if self.method == 0:
            self._op(self.A, self.p, self.ap, self.ap, 1.0, 0.0)
            nat.call("wpc_krylov_dot", self.n, self.nvec, nat.ptr(self.p), nat.ptr(self.ap),
                     self.code, pd, s)
            nat.call("wpc_krylov_update", 0, self.n, self.nvec, nat.ptr(x), nat.ptr(self.r),
                     nat.ptr(self.p), nat.ptr(self.ap), None, None, self.code, st, pd, pr, s)
            self._apply_M(self.r, self.z)
            nat.call("wpc_krylov_dot", self.n, self.nvec, nat.ptr(self.r), nat.ptr(self.z),
                     self.code, pz, s)
            nat.call("wpc_krylov_direction", 0, self.n, self.nvec, nat.ptr(self.z),
                     nat.ptr(self.p), None, None, self.code, st, pz, pr, s)
        else:
            self._apply_M(self.ap, self.map)
            nat.call("wpc_krylov_dot", self.n, self.nvec, nat.ptr(self.ap), nat.ptr(self.map),
                     self.code, pd, s)
            nat.call("wpc_krylov_update", 1, self.n, self.nvec, nat.ptr(x), nat.ptr(self.r),
                     nat.ptr(self.p), nat.ptr(self.ap), nat.ptr(self.z), nat.ptr(self.map),
                     self.code, st, pd, pr, s)
            self._op(self.A, self.z, self.az, self.az, 1.0, 0.0)
            nat.call("wpc_krylov_dot", self.n, self.nvec, nat.ptr(self.z), nat.ptr(self.az),
                     self.code, pz, s)
            nat.call("wpc_krylov_direction", 1, self.n, self.nvec, nat.ptr(self.z),
                     nat.ptr(self.p), nat.ptr(self.az), nat.ptr(self.ap), self.code, st, pz, pr, s)

    def steps(self, x, count):
        """`count` CG steps.  Where the fused step applies (see `_fused_update_step`)
        and `recurrence` is set, a step other than the last leaves its direction
        step to the next step's first operator pass (`direction_energy`), which also
        carries t = M^-1 L p in `ap` across steps; the last step ends with the
        direction step, so between calls p and the state are those of `count` calls
        of `step`."""
        pending = None
        for i in range(count):
            mode = self._recurrence_mode() if self.recurrence else None
            if mode is None:
                if pending is not None:  # (cannot happen: the mode never ends once set)
                    self._direction(pending)
                    pending = None
                self.step(x)
                continue
            last = i == count - 1
            self._recurrence_step(x, mode, pending is not None, last)
            pending = None if last else mode

    def _recurrence_mode(self):
        """("fold", count) where the fused step folds its partials in every
        workgroup, ("scalar", count) where it takes the single-workgroup folds
        (count: the operator's partials per vector), None where it does not apply
        (the preconditioner's partial count is learnt on a solve's first step)."""
        if not self.use_fused or not self.fuse_update or self._m_parts is None:
            return None
        parts_of = getattr(self.A, "cg_update_parts", None)
        count = None if parts_of is None else parts_of(self.nvec, self.dtype)
        if count is None:
            return None
        fold = max(count, self._m_parts) * self.nvec <= FOLD_IN_PLACE
        return ("fold" if fold else "scalar", count)

    def _rr_buffer(self, count):
        need = count * self.nvec
        if self._rr_parts is None or self._rr_parts.numel() < need:
            self._rr_parts = torch.zeros(need, dtype=torch.float64, device=self.r.device)
            nat.bump_generation()
        return self._rr_parts

    def _direction(self, mode):
        kind, count = mode
        st, s = nat.ptr(self.state), nat.stream()
        if kind == "fold":
            nat.call("wpc_cg_direction_fold", self.n, self.nvec, nat.ptr(self.z),
                     nat.ptr(self.p), None, self.code, self._m_parts, nat.ptr(self._m_part),
                     count, nat.ptr(self._rr_parts), st, s)
        else:  # BETA as wpc_cg_beta left it
            nat.call("wpc_cg_direction", self.n, self.nvec, nat.ptr(self.z), nat.ptr(self.p),
                     self.code, st, s)

    def _recurrence_step(self, x, mode, pending, last):
        """One fused step with t in `ap`: the first pass gathers z and takes the
        previous step's direction step when `pending`, else it is the plain first
        pass on p; the direction step runs only when `last`."""
        kind, count = mode
        rr = self._rr_buffer(count)
        st, s = nat.ptr(self.state), nat.stream()
        A = self.A
        if kind == "fold":
            if pending:
                A.direction_energy(self.z, self.p, self.ap, self.nvec, self._m_parts,
                                   self._m_part, rr, self.state)
                A.cg_update(self.p, self.r, self.nvec, self.state, rr, passes=2, x=x, t=self.ap)
            else:
                A.cg_update(self.p, self.r, self.nvec, self.state, rr, x=x, t=self.ap)
        else:
            if pending:
                A.direction_energy(self.z, self.p, self.ap, self.nvec, 0, None, None, self.state)
            else:
                A.cg_update(self.p, self.r, self.nvec, self.state, rr, passes=1, t=self.ap)
            nat.call("wpc_cg_alpha", self.nvec, count, nat.ptr(A.energy_partials(self.nvec)),
                     st, s)
            A.cg_update(self.p, self.r, self.nvec, self.state, rr, passes=6, x=x, t=self.ap)
        part, nparts = self.M.apply_batch(self.r, self.z, self.z, 1.0, 0.0, self.nvec,
                                          dot_w=self.r)
        if nparts != self._m_parts:
            raise RuntimeError("the preconditioner's partial count changed inside a run")
        self._m_part = part
        if kind == "scalar":
            nat.call("wpc_cg_beta", self.nvec, nparts, nat.ptr(part), count, nat.ptr(rr), st, s)
        if last:
            self._direction(mode)

    def _fused_update_step(self, x, s, st):
        """One CG step with the update folded into the operator's last pass
        (operators with `cg_update`: the factored squared Laplacian) -- five launches
        on the headline path instead of six: A p with the update and x += ALPHA p
        (+r.r partials) in two passes | M r (+r.z partials) in two | direction.  With partial
        counts too large to fold in every workgroup, the curvature and the r.z / r.r
        folds run as single-workgroup launches instead (the first pass | ALPHA |
        the second pass with the update and x | M r | BETA | direction: six launches
        instead of eight).  Taken once the preconditioner's partial count is known
        (from the first step); returns False to leave the step to the general path."""
        if not self.fuse_update or self._m_parts is None:
            return False
        parts_of = getattr(self.A, "cg_update_parts", None)
        if parts_of is None:
            return False
        count = parts_of(self.nvec, self.dtype)
        if count is None:
            return False
        need = count * self.nvec
        if self._rr_parts is None or self._rr_parts.numel() < need:
            self._rr_parts = torch.zeros(need, dtype=torch.float64, device=self.r.device)
            nat.bump_generation()
        if max(count, self._m_parts) * self.nvec <= FOLD_IN_PLACE:
            # (x += ALPHA p in the second pass, which reads p anyway)
            self.A.cg_update(self.p, self.r, self.nvec, self.state, self._rr_parts, x=x)
            part, nparts = self.M.apply_batch(self.r, self.z, self.z, 1.0, 0.0, self.nvec,
                                              dot_w=self.r)
            self._m_parts, self._m_part = nparts, part
            # (exact for any partial count; only slower past FOLD_IN_PLACE)
            nat.call("wpc_cg_direction_fold", self.n, self.nvec, nat.ptr(self.z),
                     nat.ptr(self.p), None, self.code, nparts, nat.ptr(part), count,
                     nat.ptr(self._rr_parts), st, s)
            return True
        self.A.cg_update(self.p, self.r, self.nvec, self.state, self._rr_parts, passes=1)
        nat.call("wpc_cg_alpha", self.nvec, count, nat.ptr(self.A.energy_partials(self.nvec)),
                 st, s)
        self.A.cg_update(self.p, self.r, self.nvec, self.state, self._rr_parts, passes=6, x=x)
        part, nparts = self.M.apply_batch(self.r, self.z, self.z, 1.0, 0.0, self.nvec,
                                          dot_w=self.r)
        self._m_parts, self._m_part = nparts, part
        nat.call("wpc_cg_beta", self.nvec, nparts, nat.ptr(part), count,
                 nat.ptr(self._rr_parts), st, s)
        nat.call("wpc_cg_direction", self.n, self.nvec, nat.ptr(self.z), nat.ptr(self.p),
                 self.code, st, s)
        return True

    def run(self, x, steps, use_graph):
        if not use_graph or steps <= 0:
            self.steps(x, steps)
            return
        # The graph bakes in x and the operators' own scratch (product partials,
        # two-pass temporaries); it is replayed only while none of them moved.
        key = (steps, x.data_ptr(), nat.generation())
        if self.graph is None or self._graph_key != key:
            # warm once eagerly (first launches load code objects), then capture
            self.graph = None
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.steps(x, steps)
            self.graph, self.graph_steps = g, steps
            # (capture itself may size a buffer: key on the generation after it)
            self._graph_key = (steps, x.data_ptr(), nat.generation())
            if self._graph_key != key:
                self.graph = None  # recorded against a buffer that has since moved
                self.steps(x, steps)
                return
        self.graph.replay()

    def host_state(self):
        s = self.state.cpu().view(-1, self.nvec)
        return s

    def snapshot_state(self, index):
        """Copy the state to pinned host memory on the current stream, without
        waiting; returns (event, host tensor).  Two buffers alternate by `index`:
        a buffer is reused only after its previous copy was waited for."""
        if getattr(self, "_snapshots", None) is None:
            self._snapshots = [torch.empty(self.state.shape, dtype=self.state.dtype,
                                           pin_memory=True) for _ in range(2)]
        host = self._snapshots[index % 2]
        host.copy_(self.state, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record()
        return ev, host


def nat_state_fields() -> int:
    return 12


def _default_tol(dtype):
    return {torch.float32: 1.0e-6, torch.float64: 1.0e-12}.get(dtype, 1.0e-3)


def _chunk_loop(work, xf, chunk, maxiter, use_graph, depth, check):
    """Run chunks of `chunk` steps (at most `maxiter` steps in all) until
    check(host_state, steps) is true for the state after some chunk.  With depth 2
    the next chunk is launched before the previous chunk's state is read (an
    asynchronous copy to pinned memory), so the device does not idle while the host
    checks; a chunk launched after every slot converged takes no steps (a resting
    slot is unchanged), so the outcome is that of checking after every chunk, at
    the cost of one chunk's launches.  Returns (steps at the deciding check, or at
    the last chunk, chunks launched)."""
    pending = []
    launched = chunks = at = 0
    while True:
        while launched < maxiter and len(pending) < depth:
            steps = min(chunk, maxiter - launched)
            work.run(xf, steps, use_graph and steps == chunk and chunks > 0)
            launched += steps
            ev, host = work.snapshot_state(chunks)
            pending.append((ev, host, launched))
            chunks += 1
        if not pending:
            break
        ev, host, at = pending.pop(0)
        ev.synchronize()
        if check(host, at):
            break
    for ev, _, _ in pending:
        ev.synchronize()
    return at, chunks


def _solve(method, A, b, x, tol, atol, maxiter, M, callback, check_every, use_cuda_graph):
    A = aslinearoperator(A)
    M = aslinearoperator(M)
    bf = b.reshape(-1)
    xf = x.reshape(-1)
    if xf.data_ptr() != x.data_ptr() or not x.is_contiguous():
        raise ValueError("x must be a contiguous device array")
    nat.require_device(bf, "b")
    if bf.dtype not in (torch.float32, torch.float64) or xf.dtype != bf.dtype:
        raise TypeError("b and x must share a float32 or float64 dtype")
    n = bf.numel()
    if maxiter == 0:
        maxiter = A.shape[0]
    if tol is None and atol is None:
        tol = _default_tol(bf.dtype)
    tol = 0.0 if tol is None else float(tol)
    atol = 0.0 if atol is None else float(atol)
    work = KrylovWork(method, A, M, n, 1, bf.dtype, bf.device)
    work.start(bf, xf, tol, atol)
    st = work.host_state()
    rr, target = float(st[1, 0]), float(st[2, 0])
    abs_tol = math.sqrt(target)
    exact = check_every == 0
    chunk = check_every if check_every > 0 else 32
    iters = 0
    if callback is not None:
        callback(0, math.sqrt(rr), abs_tol)
    if rr > target:
        def check(host, at):
            nonlocal rr, iters
            rr, iters = float(host.view(-1, 1)[1, 0]), at
            if callback is not None:
                callback(at, math.sqrt(rr), abs_tol)
            return rr <= target

        # (with a callback, which may look at x, every chunk is checked before the
        # next one starts; the count returned is that of the deciding check)
        _chunk_loop(work, xf, chunk, maxiter, use_cuda_graph, 1 if callback is not None else 2,
                    check)
        st = work.host_state()
    if exact:
        taken = int(st[4, 0])
        return (torch.tensor([taken], dtype=torch.int32, device=bf.device),
                torch.tensor([rr], dtype=torch.float64, device=bf.device), abs_tol)
    return iters, math.sqrt(rr), abs_tol


def cg(A, b, x, tol=None, atol=None, maxiter=0, M=None, callback=None, check_every=10,
       use_cuda_graph=True):
    """Preconditioned conjugate gradients; returns (iterations, residual, atol)."""
    return _solve("cg", A, b, x, tol, atol, maxiter, M, callback, check_every, use_cuda_graph)


def cr(A, b, x, tol=None, atol=None, maxiter=0, M=None, callback=None, check_every=10,
       use_cuda_graph=True):
    """Preconditioned conjugate residuals; returns (iterations, residual, atol)."""
    return _solve("cr", A, b, x, tol, atol, maxiter, M, callback, check_every, use_cuda_graph)


def solve_batched(method, A, B, X, M=None, tol=1.0e-8, atol=0.0, maxiter=10000, check_every=25,
                  use_cuda_graph=True, work=None):
    """CG/CR on k right-hand sides at once: B and X are contiguous [n, k] device arrays.

    Every product streams the matrix once for all k vectors.  Each RHS stops on
    its own target (a converged slot takes no further steps).  Returns a dict with
    per-RHS ``iterations`` (steps taken), ``residuals`` (recursive ||r||) and the
    number of ``chunks`` of ``check_every`` steps that ran.
    """
    A = aslinearoperator(A)
    M = aslinearoperator(M)
    n, k = B.shape
    if tuple(X.shape) != (n, k) or not X.is_contiguous() or not B.is_contiguous():
        raise ValueError("B and X must be contiguous [n, k] arrays")
    if work is None or work.nvec != k or work.n != n:
        work = KrylovWork(method, A, M, n, k, B.dtype, B.device)
    work.start(B.reshape(-1), X.reshape(-1), tol, atol)
    st = work.host_state()
    chunks = 0

    def converged(host, _steps):
        h = host.view(-1, k)
        return not bool((h[1] > h[2]).any())

    if not converged(st.reshape(-1), 0):
        _, chunks = _chunk_loop(work, X.reshape(-1), check_every, maxiter, use_cuda_graph, 2,
                                converged)
    st = work.host_state()
    return dict(iterations=[int(v) for v in st[4]], residuals=[float(v) ** 0.5 for v in st[1]],
                targets=[float(v) ** 0.5 for v in st[2]], chunks=chunks, work=work)


def solve_reliable(A, B, X, M, inner, tol=1.0e-8, delta=0.1, maxiter=40000, check_every=25,
                   use_cuda_graph=True, work=None, replace_every=0):
    """Mixed-precision CG with reliable updates (Sleijpen & van der Vorst, 1996): one
    float32 Krylov sequence on the operator `inner` (float32 vectors) with
    preconditioner M, whose recursive residual is replaced by the float64 true
    residual b - A x at intervals instead of restarting the inner solve.

    B, X: contiguous float64 [n, k] (X is the initial guess and the result).  The
    float32 iteration runs on the residual scaled by s_v = ||b_v - A x0_v|| (one
    scale for the whole solve) and accumulates its solution increment in float32
    Y; every `check_every` steps the recursive ||r|| is read, and when it has
    fallen below `delta` times its value at the last replacement for every moving
    right-hand side (or meets the target), or `replace_every` chunks (> 0) have
    passed since the last replacement, X += s Y in float64, Y = 0, and the
    float32 residual, its preconditioned residual and the r.z / r.r state are
    replaced from the compensated float64 residual -- the search direction p is
    kept, so the Krylov sequence continues (no restart).  A right-hand side
    finishes when its true relative residual ||b - A x|| / ||b|| <= tol (norms in
    M^-1 for the factored operator, whose mass M it is; 2-norms otherwise).

    Returns dict(iterations, replacements, relative_residual, converged, work).
    """
    from .refinement import TrueResidual

    A = aslinearoperator(A)
    inner = aslinearoperator(inner)
    M = aslinearoperator(M)
    n, k = B.shape
    if B.dtype != torch.float64 or X.dtype != torch.float64 or tuple(X.shape) != (n, k):
        raise TypeError("solve_reliable takes float64 B and X of one [n, k] shape")
    if not (B.is_contiguous() and X.is_contiguous()):
        raise ValueError("B and X must be contiguous")
    dev = B.device
    res = TrueResidual(A, None, k)  # (norms in M^-1 / M for a factored operator)
    Bf, Xf = B.reshape(-1), X.reshape(-1)
    bn = res.run(torch.zeros_like(Xf), Bf).cpu().numpy()[:k] ** 0.5
    res.run(Xf, Bf)
    scale = res.norms.clone()  # [||r0||^2 ..., ||x0||^2 ...]: the fixed residual scale
    s0 = scale.cpu().numpy()[:k] ** 0.5
    if work is None or work.nvec != k or work.n != n or work.dtype != torch.float32:
        work = KrylovWork("cg", inner, M, n, k, torch.float32, dev)
    Y = torch.zeros(n * k, dtype=torch.float32, device=dev)
    R32 = torch.empty(n * k, dtype=torch.float32, device=dev)
    full = (1 << k) - 1
    stream = nat.stream()
    # target on the scaled recursive residual: tol ||b_v|| / s_v
    target = np.array([tol * bn[v] / s0[v] if s0[v] > 0 else 0.0 for v in range(k)])
    nat.call("wpc_refine_prepare", n, k, nat.ptr(res.r), nat.ptr(scale), full, nat.ptr(R32),
             nat.code(torch.float32), stream)
    work.start(R32, Y, 0.0, 0.0)
    work._field(2).copy_(torch.tensor(target ** 2, dtype=torch.float64, device=dev))
    true_rel = np.array([res_v ** 0.5 for res_v in res.norms.cpu().numpy()[:k]]) / np.where(
        bn > 0, bn, 1.0)
    last = work.host_state()[1].numpy() ** 0.5  # recursive ||r|| at the last replacement
    replacements = 0
    steps = 0
    since = 0  # chunks since the last replacement

    def replace():
        """X += s Y; the true residual into the float32 sequence (p kept)."""
        nonlocal true_rel
        st = nat.stream()
        nat.call("wpc_refine_update", n, k, nat.ptr(Xf), nat.ptr(Y), nat.code(torch.float32),
                 nat.ptr(scale), full, None, st)
        Y.zero_()
        norms = res.run(Xf, Bf)
        nat.call("wpc_refine_prepare", n, k, nat.ptr(res.r), nat.ptr(scale), full,
                 nat.ptr(work.r), nat.code(torch.float32), st)
        work._apply_M(work.r, work.z)
        work._dot(work.r, work.z, work._field(0))
        work._dot(work.r, work.r, work._field(1))
        true_rel = norms.cpu().numpy()[:k] ** 0.5 / np.where(bn > 0, bn, 1.0)

    while steps < maxiter:
        if (true_rel <= tol).all():
            break
        chunk = min(check_every, maxiter - steps)
        work.run(Y, chunk, use_cuda_graph and chunk == check_every and steps > 0)
        steps += chunk
        h = work.host_state()
        rr = h[1].numpy() ** 0.5
        moving = true_rel > tol
        met = rr <= target
        since += 1
        if ((rr <= delta * last) | met | ~moving).all() or (replace_every and since >= replace_every):
            replace()
            replacements += 1
            since = 0
            last = work.host_state()[1].numpy() ** 0.5
    if replacements == 0 or steps >= maxiter:
        replace()
        replacements += 1
    h = work.host_state()
    return dict(iterations=[int(v) for v in h[4]], replacements=replacements,
                relative_residual=[float(v) for v in true_rel],
                converged=bool((true_rel <= tol).all()), work=work)


GMRES_MAX_RESTART = 127  # basis rows the Gram kernel takes (128 with the new vector)


def gmres(A, b, x, tol=None, atol=None, restart=31, maxiter=0, M=None, callback=None,
          check_every=31, use_cuda_graph=True, is_left_preconditioner=False):
    """Restarted right-preconditioned GMRES; returns (iterations, residual, atol).

    The signature follows warp.optim.linear.gmres as the reference's tests call it
    (tests/test_biharmonic.py:32, tests/test_indefinite.py:130,218).  Operator and
    preconditioner applies, the Gram products and basis updates (classical
    Gram-Schmidt with one reorthogonalization) and the norms run on the HIP
    kernels with the Krylov basis resident on the device; only the small
    Hessenberg least-squares problem is solved on the host, with Givens rotations
    (one device-to-host copy of the new Hessenberg column per iteration).  Left
    preconditioning is not supported.
    """
    if is_left_preconditioner:
        raise NotImplementedError("gmres supports right preconditioning only")
    A = aslinearoperator(A)
    M = aslinearoperator(M)
    bf, xf = b.reshape(-1), x.reshape(-1)
    nat.require_device(bf, "b")
    n = bf.numel()
    dt = bf.dtype
    code = nat.code(dt)
    if maxiter == 0:
        maxiter = A.shape[0]
    if tol is None and atol is None:
        tol = _default_tol(dt)
    tol = 0.0 if tol is None else float(tol)
    atol = 0.0 if atol is None else float(atol)
    m = max(1, min(int(restart), GMRES_MAX_RESTART))
    dev = bf.device
    V = torch.empty((m + 1, n), dtype=dt, device=dev)
    Z = torch.empty((m, n), dtype=dt, device=dev) if M is not None else V
    r = torch.empty_like(bf)
    w = torch.empty_like(bf)
    s = nat.stream()
    parts = torch.empty(nat.lib().wpc_gram_partials(n) * (m + 1), dtype=torch.float64, device=dev)
    dparts = torch.empty(nat.lib().wpc_dot_partials(n), dtype=torch.float64, device=dev)
    # per iteration: h (m+1), h2 (m+1), ||w||^2 (1); the final coefficients y (m)
    col = torch.zeros(2 * (m + 1) + 1, dtype=torch.float64, device=dev)
    ycoef = torch.empty(m, dtype=torch.float64, device=dev)

    def norm(v):
        nat.call("wpc_dot", n, 1, nat.ptr(v), nat.ptr(v), code, nat.ptr(dparts), nat.ptr(col[-1:]),
                 s)
        return math.sqrt(max(float(col[-1].item()), 0.0))

    def scaled(src, dst, factor):
        nat.call("wpc_vector_scale", n, 1, nat.ptr(src), nat.ptr(dst), float(factor), 0.0, 0.0,
                 code, s)

    def residual():
        A.matvec(xf, bf, r, -1.0, 1.0)
        return norm(r)

    bnorm = norm(bf)
    target = max(atol, tol * bnorm)
    beta = residual()
    iters = 0
    if callback is not None:
        callback(0, beta, target)
    while iters < maxiter and beta > target:
        scaled(r, V[0], 1.0 / beta)
        H = np.zeros((m + 1, m))
        cs, sn = np.zeros(m), np.zeros(m)
        g = np.zeros(m + 1)
        g[0] = beta
        k = 0
        for j in range(m):
            if M is not None:
                M.matvec(V[j], Z[j], Z[j], 1.0, 0.0)
            A.matvec(Z[j], w, w, 1.0, 0.0)
            h, h2 = col[: j + 1], col[m + 1: m + 2 + j]
            for coef in (h, h2):
                nat.call("wpc_gram", n, j + 1, nat.ptr(V), n, nat.ptr(w), code, nat.ptr(parts),
                         nat.ptr(coef), s)
                nat.call("wpc_basis_combine", n, j + 1, nat.ptr(V), n, nat.ptr(coef), -1.0,
                         nat.ptr(w), code, s)
            nat.call("wpc_dot", n, 1, nat.ptr(w), nat.ptr(w), code, nat.ptr(dparts),
                     nat.ptr(col[-1:]), s)
            host = col.cpu().numpy()
            hcol = host[: j + 1] + host[m + 1: m + 2 + j]
            hn = math.sqrt(max(float(host[-1]), 0.0))
            H[: j + 1, j] = hcol
            H[j + 1, j] = hn
            if hn > 0:
                scaled(w, V[j + 1], 1.0 / hn)
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            d = math.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = (1.0, 0.0) if d == 0 else (H[j, j] / d, H[j + 1, j] / d)
            H[j, j] = d
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            k = j + 1
            iters += 1
            if abs(g[j + 1]) <= target or iters >= maxiter or hn == 0:
                break
        y = np.zeros(k)
        for i in range(k - 1, -1, -1):
            y[i] = (g[i] - H[i, i + 1:k] @ y[i + 1:k]) / H[i, i]
        ycoef[:k].copy_(torch.from_numpy(y))
        nat.call("wpc_basis_combine", n, k, nat.ptr(Z), n, nat.ptr(ycoef), 1.0, nat.ptr(xf), code,
                 s)
        beta = residual()
        if callback is not None:
            callback(iters, beta, target)
    return iters, beta, target


def solve_refined(method, A, B, X, M=None, inner=None, tol=1.0e-8, inner_tol=1.0e-4,
                  max_outer=20, maxiter=100000, check_every=25):
    """Iterative refinement to a true relative residual ``tol`` on k right-hand sides.

    ``A`` is the float64 system as a ``SparseOperator``; its residual b - A x is
    recomputed with compensated accumulation (wpc_dd_residual), so the reported
    residual is not limited by float64 cancellation (relevant when eps * cond(A)
    approaches ``tol``).  Each outer step solves ``inner`` (default A; e.g. a
    float32 copy with a matching preconditioner M) for the correction of the
    still-failing right-hand sides, scaled to unit size, with CG/CR to
    ``inner_tol`` and adds it in float64.  B and X are float64 [n, k] (or [n])
    device arrays; X is the initial guess and receives the solution.  Returns
    per-RHS ``iterations`` (inner steps summed), ``relative_residual`` and
    ``outer`` (steps taken).
    """
    from .refinement import TrueResidual

    squeeze = B.dim() == 1
    B2 = B.reshape(B.shape[0], -1)
    X2 = X.reshape(X.shape[0], -1)
    n, k = B2.shape
    if B2.dtype != torch.float64 or X2.dtype != torch.float64:
        raise TypeError("solve_refined takes float64 B and X")
    if not (X2.is_contiguous() and B2.is_contiguous()):
        raise ValueError("B and X must be contiguous")
    inner = A if inner is None else inner
    vdt = getattr(inner, "scalar_type", torch.float64)
    dev = B2.device
    res = TrueResidual(A, None, k)
    zero = torch.zeros_like(X2)
    bn = torch.sqrt(res.run(zero.reshape(-1), B2.reshape(-1))[:k]).cpu().numpy()
    norms = res.run(X2.reshape(-1), B2.reshape(-1)).cpu().numpy()
    Rin = torch.empty(n * k, dtype=vdt, device=dev)
    C = torch.empty(n * k, dtype=vdt, device=dev)
    iters = [0] * k
    work = None
    outer = 0
    scale = lambda v: v if v > 0 else 1.0  # noqa: E731
    while outer < max_outer:
        rel = [float(norms[v]) ** 0.5 / scale(bn[v]) for v in range(k)]
        fail = [rel[v] > tol for v in range(k)]
        if not any(fail):
            break
        mask = sum(1 << v for v in range(k) if fail[v])
        s = nat.stream()
        nat.call("wpc_refine_prepare", n, k, nat.ptr(res.r), nat.ptr(res.norms), mask, nat.ptr(Rin),
                 nat.code(vdt), s)
        C.zero_()
        out = solve_batched(method, inner, Rin.view(n, k), C.view(n, k), M=M, tol=inner_tol,
                            maxiter=maxiter, check_every=check_every, work=work)
        work = out["work"]
        iters = [a + b for a, b in zip(iters, out["iterations"])]
        nat.call("wpc_refine_update", n, k, nat.ptr(X2), nat.ptr(C), nat.code(vdt),
                 nat.ptr(res.norms), mask, None, nat.stream())
        norms = res.run(X2.reshape(-1), B2.reshape(-1)).cpu().numpy()
        outer += 1
    rel = [float(norms[v]) ** 0.5 / scale(bn[v]) for v in range(k)]
    out = dict(iterations=iters, relative_residual=rel, outer=outer,
               converged=all(r <= tol for r in rel))
    if squeeze:
        out["iterations"], out["relative_residual"] = out["iterations"][0], out["relative_residual"][0]
    return out
