"""Device assembly of the operators: P1 stiffness/mass on tetrahedra and the
assembled squared operator Q = alpha*M + K M^-1 K (HIP kernels; torch only sorts
keys and allocates).

The reference assembles these on the host (benchmarks/elasticity_problem.py:33-61,
visualization/solve_dirichlet.py:150-153); the serving loop needs them on the device
so that setup latency does not scale with host SpGEMM time.  ``squared_operator``
reproduces the host product's term order (descending k along each row, the order
scipy's intermediate K @ diag(1/m) stores), so for the same K every entry is
bit-identical to ``(K @ diag(1/m)) @ K + alpha diag(m)`` formed with sequential
accumulation.  Its pattern is structural (every column reachable through a stored
entry of K), so it can hold exact zeros that a value-dropping host product omits;
a fixed pattern is what factor reuse across value updates needs.
"""

from __future__ import annotations

import itertools

import numpy as np
import torch

from . import _native as nat
from .sparse import BsrMatrix


def _csr_from_sorted_keys(nrow, ncol, keys, values, device):
    rows = keys // ncol
    cols = (keys % ncol).to(torch.int32)
    counts = torch.bincount(rows, minlength=nrow)
    offsets = torch.zeros(nrow + 1, dtype=torch.int32, device=device)
    offsets[1:] = torch.cumsum(counts, 0).to(torch.int32)
    return BsrMatrix(nrow, ncol, (1, 1), offsets, cols.contiguous(), values)


def assemble_triplets(nrow, ncol, rows, cols, values):
    """Scalar CSR from triplets, duplicates summed in input order (deterministic)."""
    device = values.device
    key = rows.to(torch.int64) * ncol + cols.to(torch.int64)
    order = torch.argsort(key, stable=True)
    skey = key[order]
    uniq, counts = torch.unique_consecutive(skey, return_counts=True)
    seg = torch.zeros(uniq.numel() + 1, dtype=torch.int64, device=device)
    seg[1:] = torch.cumsum(counts, 0)
    out = torch.empty(uniq.numel(), dtype=torch.float64, device=device)
    nat.call("wpc_segment_sum", uniq.numel(), nat.ptr(seg), nat.ptr(order), nat.ptr(values),
             nat.ptr(out), nat.stream())
    return _csr_from_sorted_keys(nrow, ncol, uniq, out, device)


def kuhn_tets(shape, device):
    """Tetrahedra (int32 [T, 4]) of the 6-tet Kuhn split of a grid of vertex shape
    (vertex id (x*ny + y)*nz + z), in the order of workloads.kuhn_grid."""
    nx, ny, nz = shape
    base = torch.stack(torch.meshgrid(torch.arange(nx - 1, device=device),
                                      torch.arange(ny - 1, device=device),
                                      torch.arange(nz - 1, device=device), indexing="ij"),
                       dim=-1).reshape(-1, 3)
    cells = []
    for order in itertools.permutations(range(3)):
        corners = [np.zeros(3, dtype=np.int64)]
        for axis in order:
            c = corners[-1].copy()
            c[axis] += 1
            corners.append(c)
        cidx = base[:, None, :] + torch.tensor(np.array(corners), device=device)[None]
        cells.append((cidx[..., 0] * ny + cidx[..., 1]) * nz + cidx[..., 2])
    return torch.cat(cells).to(torch.int32).contiguous()


def p1_laplacian(vertices: torch.Tensor, tets: torch.Tensor):
    """P1 stiffness K (BsrMatrix, float64) and lumped mass (float64) on the device."""
    nv, nt = vertices.shape[0], tets.shape[0]
    dev = vertices.device
    local = torch.empty(nt * 16, dtype=torch.float64, device=dev)
    quarter = torch.empty(nt, dtype=torch.float64, device=dev)
    nat.call("wpc_p1_local", nt, nat.ptr(vertices.contiguous()), nat.ptr(tets.contiguous()),
             nat.ptr(local), nat.ptr(quarter), nat.stream())
    rows = tets.repeat_interleave(4, dim=1).reshape(-1)
    cols = tets.repeat(1, 4).reshape(-1)
    K = assemble_triplets(nv, nv, rows, cols, local)
    vid = tets.reshape(-1).to(torch.int64)
    order = torch.argsort(vid, stable=True)
    counts = torch.bincount(vid, minlength=nv)
    seg = torch.zeros(nv + 1, dtype=torch.int64, device=dev)
    seg[1:] = torch.cumsum(counts, 0)
    mass = torch.empty(nv, dtype=torch.float64, device=dev)
    shares = quarter.repeat_interleave(4)
    nat.call("wpc_segment_sum", nv, nat.ptr(seg), nat.ptr(order), nat.ptr(shares),
             nat.ptr(mass), nat.stream())
    return K, mass


def squared_operator(K: BsrMatrix, mass: torch.Tensor, alpha: float = 1.0) -> BsrMatrix:
    """Q = K diag(1/mass) K + alpha diag(mass) (float64) by the device SpGEMM."""
    n = K.nrow
    dev = K.device
    s = nat.stream()
    inverse = torch.empty_like(mass)
    status = torch.zeros(3, dtype=torch.int32, device=dev)
    nat.call("wpc_sqlap_mass_inverse", n, nat.ptr(mass), nat.ptr(inverse), nat.code(mass.dtype),
             nat.ptr(status), s)
    counts = torch.empty(n, dtype=torch.int32, device=dev)
    nat.call("wpc_spgemm_kmk_count", n, nat.ptr(K.offsets), nat.ptr(K.columns),
             nat.ptr(counts), s)
    if int(status[2].item()):
        raise ValueError("mass entries must be finite and positive")
    if bool((counts < 0).any().item()):
        raise ValueError("a row of K M^-1 K exceeds 512 columns")
    offsets = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    nat.call("wpc_scan_inclusive", n, nat.ptr(counts), nat.ptr(offsets[1:]), s)
    nnz = int(offsets[-1].item())
    cols = torch.empty(nnz, dtype=torch.int32, device=dev)
    vals = torch.empty(nnz, dtype=torch.float64, device=dev)
    nat.call("wpc_spgemm_kmk_fill", n, nat.ptr(K.offsets), nat.ptr(K.columns), nat.ptr(K.values),
             nat.ptr(inverse), nat.ptr(mass), float(alpha), nat.ptr(offsets), nat.ptr(cols),
             nat.ptr(vals), s)
    return BsrMatrix(n, n, (1, 1), offsets, cols, vals)


def add_diagonal(A: BsrMatrix, d: torch.Tensor, alpha: float) -> BsrMatrix:
    """A + alpha*diag(d) as a new matrix sharing A's index buffers (float64)."""
    values = A.values.clone()
    status = torch.zeros(1, dtype=torch.int32, device=A.device)
    nat.call("wpc_add_diagonal", A.nrow, nat.ptr(A.offsets), nat.ptr(A.columns), nat.ptr(values),
             nat.ptr(d), float(alpha), nat.ptr(status), nat.stream())
    if int(status.item()):
        raise ValueError("add_diagonal requires a stored diagonal in every row")
    return BsrMatrix(A.nrow, A.ncol, (1, 1), A.offsets, A.columns, values)


def jittered_grid(side: int, seed: int, device, jitter: float = 0.15):
    """Vertices (float64 [side^3, 3]) of the unit-cube grid with interior vertices
    moved uniformly by +-jitter*h (a device generator seeded with `seed`), and the
    Kuhn tetrahedra of that grid."""
    h = 1.0 / (side - 1)
    ax = torch.arange(side, device=device)
    idx = torch.stack(torch.meshgrid(ax, ax, ax, indexing="ij"), dim=-1).reshape(-1, 3)
    gen = torch.Generator(device=device)
    gen.manual_seed(int(seed))
    noise = (torch.rand(idx.shape, generator=gen, dtype=torch.float64, device=device) * 2 - 1) * (
        jitter * h)
    interior = ((idx > 0) & (idx < side - 1)).all(dim=1, keepdim=True)
    vertices = idx.to(torch.float64) * h + torch.where(interior, noise, torch.zeros_like(noise))
    return vertices.contiguous(), kuhn_tets((side, side, side), device)
