"""Fixed-support FSAI refits and the quality estimate (drop-in for fsai_reuse.py).

``RefitPlan`` snapshots the source topology, maps every scalar nonzero to its
position in the source BSR values and every G nonzero to its position in G^T
(fsai_reuse.py:157-184).  ``update`` checks the topology on the device, gathers
the new values, equilibrates, refits each row's stored support with a dense
Cholesky, validates, and only then commits into the existing G / G^T buffers
(fsai_reuse.py:195-245), so captured applies see the new factor and a failed
refit leaves the previous one intact.
"""

from __future__ import annotations

import math

import torch

from . import _native as nat
from .linear import aslinearoperator
from .sparse import BsrMatrix


class RefitPlan:
    def __init__(self, owner, A: BsrMatrix, scalar: BsrMatrix, width: int, floor: float):
        self.signature = (A.shape, A.dtype, A.device, A.row_counts is not None)
        self.topology = [t.clone() for t in self.arrays(A)]
        self.scalar = scalar
        self.floor = floor
        self.width = width
        dev = A.device
        s = nat.stream()
        self.mapping = torch.empty(scalar.nnz, dtype=torch.int32, device=dev)
        nat.call("wpc_refit_source_map", scalar.nrow, nat.ptr(scalar.offsets),
                 nat.ptr(scalar.columns), nat.ptr(A.offsets), nat.ptr(A.columns),
                 A.block_shape[0], nat.ptr(self.mapping), s)
        G, GT = owner.G, owner.GT
        self.transpose = torch.empty(G.nnz, dtype=torch.int32, device=dev)
        nat.call("wpc_refit_transpose_map", G.nrow, nat.ptr(G.offsets), nat.ptr(G.columns),
                 nat.ptr(GT.offsets), nat.ptr(GT.columns), nat.ptr(self.transpose), s)
        self.values = torch.empty_like(scalar.values)
        self.candidate = torch.empty_like(G.values)
        self.scale = torch.empty(scalar.nrow, dtype=A.scalar_type, device=dev)
        self.bad = torch.zeros(3, dtype=torch.int32, device=dev)
        self.scaled = owner.row_scale2 is not None
        if self.scaled:
            self.exact = torch.empty(G.nnz, dtype=A.scalar_type, device=dev)
            self.candidate_scale2 = torch.empty_like(owner.row_scale2)

    @staticmethod
    def arrays(A: BsrMatrix):
        return [A.offsets[: A.nrow + 1], A.columns[: A.nnz]]

    def update(self, owner, A: BsrMatrix):
        if (A.shape, A.dtype, A.device, A.row_counts is not None) != self.signature:
            raise ValueError("FSAI update requires the original shape, dtype, device and topology")
        incoming = self.arrays(A)
        if any(a.shape != b.shape for a, b in zip(incoming, self.topology)):
            raise ValueError("FSAI update requires identical topology storage")
        s = nat.stream()
        self.bad.zero_()
        for a, b in zip(incoming, self.topology):
            nat.call("wpc_compare_i32", a.numel(), nat.ptr(a), nat.ptr(b), nat.ptr(self.bad), s)
        if int(self.bad[0].item()):
            raise ValueError("FSAI update requires identical sparsity; rebuild to change it")
        if A.values.numel() < self.topology[1].numel():
            raise ValueError("Insufficient source values")
        sc = self.scalar
        dt = nat.code(A.scalar_type)
        nat.call("wpc_gather", self.mapping.numel(), A.values.element_size(), nat.ptr(A.values),
                 nat.ptr(self.mapping), nat.ptr(self.values), s)
        nat.call("wpc_fsai_diagonal", sc.nrow, nat.ptr(sc.offsets), nat.ptr(sc.columns),
                 nat.ptr(self.values), dt, nat.ptr(self.scale), nat.ptr(self.bad), s)
        G, GT = owner.G, owner.GT
        out = self.exact if self.scaled else self.candidate
        out_dt = A.scalar_type if self.scaled else G.scalar_type
        nat.call("wpc_fsai_refit", sc.nrow, nat.ptr(sc.offsets), nat.ptr(sc.columns),
                 nat.ptr(self.values), dt, nat.ptr(G.offsets), nat.ptr(G.columns),
                 nat.ptr(self.scale), float(self.floor), self.width, nat.ptr(out),
                 nat.code(out_dt), nat.ptr(self.bad), s)
        if self.scaled:
            nat.call("wpc_fsai_compress_scaled", G.nrow, nat.ptr(G.offsets), nat.ptr(G.columns),
                     nat.ptr(self.exact), dt, nat.ptr(self.candidate), nat.code(G.scalar_type),
                     nat.ptr(self.candidate_scale2), nat.ptr(self.bad), s)
        if int(self.bad[0].item()):
            raise ValueError("FSAI refit requires finite positive diagonals and SPD local supports")
        G.values.copy_(self.candidate)
        if self.scaled:
            owner.row_scale2.copy_(self.candidate_scale2)
        nat.call("wpc_scatter", self.transpose.numel(), G.values.element_size(),
                 nat.ptr(G.values), nat.ptr(self.transpose), nat.ptr(GT.values), s)
        if owner._plans is not None:
            for plan in owner._plans:
                plan.refresh_values()


def quality(owner, A, probes=4, seed=17):
    """Rademacher estimate of ||G A G^T - I||_F / sqrt(n) (fsai_reuse.py:248-279).

    The probe stream is this library's own (splitmix64), not Warp's; the estimate
    is a diagnostic, exact only in expectation.
    """
    if not isinstance(probes, int) or probes < 1:
        raise ValueError("probes must be a positive integer")
    if (A.shape, A.dtype, A.device) != (owner.shape, owner.dtype, owner.device):
        raise ValueError("quality requires matching shape, dtype and device")
    from .sparse_operator import _matvec

    op = aslinearoperator(A)
    n = A.shape[0]
    dt = A.scalar_type
    x, y, z, w = (torch.empty(n, dtype=dt, device=A.device) for _ in range(4))
    scale = owner.row_scale2
    parts = torch.empty(nat.lib().wpc_dot_partials(n), dtype=torch.float64, device=A.device)
    acc = torch.empty(1, dtype=torch.float64, device=A.device)
    total = 0.0
    s = nat.stream()
    for p in range(probes):
        nat.call("wpc_rademacher", n, seed + p, nat.ptr(x), nat.code(dt), s)
        if scale is None:
            _matvec(owner.GT, x, y, y, 1.0, 0.0, owner.apply_lanes)
            op.matvec(y, z, z, 1.0, 0.0)
            _matvec(owner.G, z, w, w, 1.0, 0.0, owner.apply_lanes)
        else:  # G = diag(s) q: G^T x = q^T (s x)
            sx = x * scale.to(dt).sqrt()
            _matvec(owner.GT, sx, y, y, 1.0, 0.0, owner.apply_lanes)
            op.matvec(y, z, z, 1.0, 0.0)
            _matvec(owner.G, z, w, w, 1.0, 0.0, owner.apply_lanes)
            w.mul_(scale.to(dt).sqrt())
        nat.call("wpc_axpby", n, 1.0, nat.ptr(w), -1.0, nat.ptr(x), nat.ptr(w), nat.code(dt), s)
        nat.call("wpc_dot", n, 1, nat.ptr(w), nat.ptr(w), nat.code(dt), nat.ptr(parts),
                 nat.ptr(acc), s)
        total += float(acc.item())
    return math.sqrt(total / (probes * max(n, 1)))


def refit_clone(template, A: BsrMatrix, topology=None):
    """A new FSAI for A on `template`'s supports: the reuse path of
    fsai_reuse.py:195-245 applied to a second matrix of the same topology.

    The clone shares the template's index buffers (G/GT offsets and columns, row
    partitions) and owns its values, so the template keeps serving its own
    matrix.  ``topology`` = (offsets, columns) of the template's source (default:
    the template's source matrix).  Synchronizes once to read the refit status.
    """
    from .fsai import FSAI
    from .sparse import StreamPlan

    if not isinstance(A, BsrMatrix) or A.block_shape != (1, 1) or A.row_counts is not None:
        raise ValueError("refit_clone requires a compact scalar CSR matrix")
    n = template._scalar_n
    if A.shape != (n, n):
        raise ValueError("FSAI update requires the original shape, dtype, device and topology")
    off0, col0 = topology if topology is not None else RefitPlan.arrays(template.source)
    inc = RefitPlan.arrays(A)
    if inc[0].shape != off0.shape or inc[1].shape != col0.shape:
        raise ValueError("FSAI update requires identical topology storage")
    dev = A.device
    s = nat.stream()
    bad = torch.zeros(3, dtype=torch.int32, device=dev)
    nat.call("wpc_compare_i32", off0.numel(), nat.ptr(inc[0]), nat.ptr(off0), nat.ptr(bad), s)
    nat.call("wpc_compare_i32", col0.numel(), nat.ptr(inc[1]), nat.ptr(col0), nat.ptr(bad), s)
    G, GT = template.G, template.GT
    dt = nat.code(A.scalar_type)
    scale = torch.empty(n, dtype=A.scalar_type, device=dev)
    nat.call("wpc_fsai_diagonal", n, nat.ptr(A.offsets), nat.ptr(A.columns), nat.ptr(A.values), dt,
             nat.ptr(scale), nat.ptr(bad[1:]), s)
    scaled = template.row_scale2 is not None
    values = torch.empty(G.nnz, dtype=G.scalar_type, device=dev)
    exact = torch.empty(G.nnz, dtype=A.scalar_type, device=dev) if scaled else values
    nat.call("wpc_fsai_refit", n, nat.ptr(A.offsets), nat.ptr(A.columns), nat.ptr(A.values), dt,
             nat.ptr(G.offsets), nat.ptr(G.columns), nat.ptr(scale), float(template.pivot_floor),
             template.max_row_size, nat.ptr(exact), nat.code(exact.dtype), nat.ptr(bad[1:]), s)
    row_scale2 = None
    if scaled:
        row_scale2 = torch.empty_like(template.row_scale2)
        nat.call("wpc_fsai_compress_scaled", n, nat.ptr(G.offsets), nat.ptr(G.columns),
                 nat.ptr(exact), dt, nat.ptr(values), nat.code(G.scalar_type), nat.ptr(row_scale2),
                 nat.ptr(bad[1:]), s)
    status = bad.cpu()
    if int(status[0]):
        raise ValueError("FSAI update requires identical sparsity; rebuild to change it")
    if int(status[1]):
        raise ValueError("FSAI refit requires finite positive diagonals and SPD local supports")
    transpose = getattr(template, "_transpose_map", None)
    if transpose is None:
        transpose = torch.empty(G.nnz, dtype=torch.int32, device=dev)
        nat.call("wpc_refit_transpose_map", n, nat.ptr(G.offsets), nat.ptr(G.columns),
                 nat.ptr(GT.offsets), nat.ptr(GT.columns), nat.ptr(transpose), s)
        template._transpose_map = transpose
    gt_values = torch.empty_like(values)
    nat.call("wpc_scatter", G.nnz, values.element_size(), nat.ptr(values), nat.ptr(transpose),
             nat.ptr(gt_values), s)
    clone = FSAI.__new__(FSAI)
    for name in ("max_row_size", "max_step_size", "factor_dtype", "apply_lanes", "kap_tolerance",
                 "pivot_floor", "truncated_rows", "_scalar_n"):
        setattr(clone, name, getattr(template, name))
    clone.source = A
    clone.G = BsrMatrix(n, n, (1, 1), G.offsets, G.columns, values)
    clone.GT = BsrMatrix(n, n, (1, 1), GT.offsets, GT.columns, gt_values)
    clone.row_scale2 = row_scale2
    clone._transpose_map = transpose
    clone._tmp = {}
    clone._refit = None
    clone._plans = None
    if template._plans is not None:
        clone._plans = (StreamPlan.like(template._plans[0], clone.G),
                        StreamPlan.like(template._plans[1], clone.GT))
    LinearOperator_init(clone, A)
    return clone


def LinearOperator_init(op, A):
    from .linear import LinearOperator

    LinearOperator.__init__(op, A.shape, A.dtype, A.device, op._apply)
