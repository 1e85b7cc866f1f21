"""Block-adaptive FSAI on b x b BSR matrices (the reference's experimental
benchmarks/block_fsai_candidate.py: BlockFSAI and BsrBlockFSAI).

Each block row grows a block support greedily (largest block residual score in
the strictly lower block frontier) and is normalized so that G_i A G_i^T = I.  The
factor is stored as scalar CSR (G, G^T) and applied with the same streamed/sliced
products as the scalar FSAI (so it also drives the fused CG path);
``BsrBlockFSAI`` additionally keeps b x b BSR copies and applies through the BSR
product, as the reference's variant does.
"""

from __future__ import annotations

import torch

from . import _native as nat
from .fsai import FSAI, _pack_factor
from .sparse import BsrMatrix, bsr_copy, bsr_transposed


class BlockFSAI(FSAI):
    def __init__(self, a, max_blocks=2, apply_lanes=1, factor_dtype=torch.float32,
                 reuse_pattern=False):
        if not isinstance(a, BsrMatrix):
            raise ValueError("BlockFSAI requires a Warp BSR matrix")
        if a.row_counts is not None:
            a = bsr_copy(a)
        b = a.block_shape[0]
        if b != a.block_shape[1] or b < 2:
            raise ValueError("BlockFSAI requires square blocks of size 2 or more")
        if a.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("BlockFSAI supports float32 and float64")
        nat.require_device(a.values, "A")
        n = a.shape[0]
        dtype, device = a.scalar_type, a.device
        width = b * max_blocks
        scale = torch.empty(n, dtype=dtype, device=device)
        status = torch.zeros(1, dtype=torch.int32, device=device)
        cols = torch.empty(n * width, dtype=torch.int32, device=device)
        vals = torch.empty(n * width, dtype=dtype, device=device)
        self._build(a, b, max_blocks, scale, cols, vals, status, refit=False)
        self.max_blocks = max_blocks
        self.max_row_size = width
        self.max_step_size = 1
        self.kap_tolerance = 0.0
        self.pivot_floor = 0.0
        self.truncated_rows = 0
        self.factor_dtype = factor_dtype
        self.apply_lanes = apply_lanes
        self.row_scale2 = None
        self.source = a
        self._publish(n, width, cols, vals, dtype)
        self._tmp = {}
        from .linear import LinearOperator

        LinearOperator.__init__(self, a.shape, a.dtype, device, self._apply)
        self._scalar_n = n
        self._refit = None
        self._reuse = (a, b, width, scale, status, cols, vals) if reuse_pattern else None

    @staticmethod
    def _build(a, b, max_blocks, scale, cols, vals, status, refit):
        nat.call("wpc_block_fsai_build", a.nrow, b, max_blocks, nat.ptr(a.offsets),
                 nat.ptr(a.columns), nat.ptr(a.values), nat.code(a.scalar_type), nat.ptr(scale),
                 nat.ptr(cols), nat.ptr(vals), nat.ptr(status), int(refit), nat.stream())
        if int(status.item()):
            raise ValueError("Nonpositive block FSAI pivot")

    def _publish(self, n, width, cols, vals, dtype):
        # _pack_factor sorts the padded rows in place (as fsai_pack.py does); a refit
        # reads its block pattern back from them
        G = _pack_factor(n, width, cols, vals)
        if self.factor_dtype != dtype:
            G = bsr_copy(G, scalar_type=self.factor_dtype)
        self.G = G
        self.GT = bsr_transposed(G)
        self._plans = None

    def update(self, a=None):
        """Fixed-support refit (block_fsai_candidate.py:217-254): identical topology
        required; replaces the factor buffers (recapture graphs afterwards); a failed
        update leaves the published factor unchanged."""
        if self._reuse is None:
            raise ValueError("Construct with reuse_pattern=True")
        old, b, width, scale, status, cols, vals = self._reuse
        if a is None:
            a = old
        if (a.shape, a.dtype, a.device) != (old.shape, old.dtype, old.device):
            raise ValueError("refit requires identical shape, dtype and device")
        if a.row_counts is not None:
            a = bsr_copy(a)
        if a.nnz_sync() != old.nnz_sync():
            raise ValueError("refit requires identical topology")
        status.zero_()
        s = nat.stream()
        nat.call("wpc_compare_i32", a.nrow + 1, nat.ptr(a.offsets), nat.ptr(old.offsets),
                 nat.ptr(status), s)
        nat.call("wpc_compare_i32", a.nnz, nat.ptr(a.columns), nat.ptr(old.columns),
                 nat.ptr(status), s)
        if int(status.item()):
            raise ValueError("refit requires identical topology")
        new_cols, new_vals = cols.clone(), torch.empty_like(vals)
        self._build(a, b, self.max_blocks, scale, new_cols, new_vals, status, refit=True)
        self._publish(a.shape[0], width, new_cols, new_vals, a.scalar_type)
        self._reuse = (a, b, width, scale, status, new_cols, new_vals)
        self.source = a
        return self


def _blockify(C: BsrMatrix, b: int) -> BsrMatrix:
    """b x b BSR copy of a scalar CSR matrix whose b scalar rows per block row hold
    the same block columns with every component stored (a block FSAI factor)."""
    dev = C.device
    nb = C.nrow // b
    off = C.offsets.long()
    boff = (off[::b] // (b * b)).to(torch.int32).contiguous()  # b*b scalar entries per block
    nnzb = int(boff[-1].item())
    counts = (boff[1:] - boff[:-1]).long()
    brow = torch.repeat_interleave(torch.arange(nb, device=dev), counts)
    k = torch.arange(nnzb, device=dev) - boff[:-1].long()[brow]
    first = off[brow * b] + k * b  # first scalar entry of block (brow, k) in its row 0
    bcol = (C.columns.long()[first] // b).to(torch.int32)
    r = torch.arange(b, device=dev)
    src = off[brow[:, None] * b + r[None, :]] + (k * b)[:, None]  # [nnzb, b] row starts
    idx = src[:, :, None] + r[None, None, :]  # [nnzb, b(r), b(c)]
    vals = C.values[idx.reshape(-1)].reshape(nnzb, b, b).contiguous()
    return BsrMatrix(nb, C.ncol // b, (b, b), boff, bcol.contiguous(), vals)


class BsrBlockFSAI(BlockFSAI):
    """BlockFSAI applied through b x b block products (block_fsai_candidate.py:292-332)."""

    def _publish(self, n, width, cols, vals, dtype):
        super()._publish(n, width, cols, vals, dtype)
        b = self.source.block_shape[0] if hasattr(self, "source") else width // self.max_blocks
        self.B = _blockify(self.G, b)
        self.BT = _blockify(self.GT, b)

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        if nvec != 1 or dot_w is not None:
            return super().apply_batch(x, y, z, alpha, beta, nvec, dot_w, dot_out)
        self._apply(x, y, z, alpha, beta)

    fused = False

    def _apply(self, x, y, z, alpha, beta):
        x, y, z = x.reshape(-1), y.reshape(-1), z.reshape(-1)
        if x.dtype != self.B.scalar_type:
            return FSAI._apply(self, x, y, z, alpha, beta)
        b = self.B.block_shape[0]
        tmp = self._scratch(x.dtype, self._scalar_n)
        s = nat.stream()
        nat.call("wpc_bsr_spmv", self.B.nrow, b, nat.ptr(self.B.offsets), nat.ptr(self.B.columns),
                 nat.ptr(self.B.values), nat.code(self.B.scalar_type), nat.ptr(x), None,
                 nat.ptr(tmp), 1.0, 0.0, s)
        if z.data_ptr() == tmp.data_ptr():
            raise ValueError("z must not alias the scratch")
        nat.call("wpc_bsr_spmv", self.BT.nrow, b, nat.ptr(self.BT.offsets),
                 nat.ptr(self.BT.columns), nat.ptr(self.BT.values), nat.code(self.BT.scalar_type),
                 nat.ptr(tmp), nat.ptr(y), nat.ptr(z), float(alpha), float(beta), s)
