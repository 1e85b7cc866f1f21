"""Matrix Market ingest of the reference's benchmark dumps.

``load_dump(directory, k)`` reads the files the reference's dragon benchmark
reads (benchmarks/dragon.py:79-96): ``k{k}_Q.mtx`` (the order-k system) and
``k{k}_rhs.mtx``; for k = 2 also ``k1_Q.mtx`` (H = M + K) and ``k4_Q.mtx`` (the
mixed system, whose leading diagonal is the lumped mass and whose upper-right
block is -K), checks H = M + K and Q = M + K M^-1 K as the benchmark does, and
uploads everything as device matrices.  Host parsing is scipy's; the data
themselves are not shipped (the benchmark downloads them separately).
"""

from __future__ import annotations

from pathlib import Path

import numpy as np
import scipy.io as sio
import scipy.sparse as ss
import torch

from .sparse import BsrMatrix, from_scipy


def upload(a, device=None, dtype=torch.float64) -> BsrMatrix:
    """Scalar CSR device matrix of a scipy sparse matrix (duplicates summed)."""
    a = ss.csr_matrix(a)
    a.sum_duplicates()
    a.sort_indices()
    device = torch.device("cuda", 0) if device is None else torch.device(device)
    return from_scipy(a, device, dtype)


def backward_error(a, abs_a, b, x) -> float:
    """Componentwise backward error max_i |b - A x|_i / (|A| |x| + |b|)_i
    (benchmarks/dragon.py:34-37), on the host."""
    r = np.abs(b - a @ x)
    scale = abs_a @ np.abs(x) + np.abs(b)
    return float(np.max(r / np.maximum(scale, np.finfo(float).tiny)))


def load_dump(directory, k=2, device=None, dtype=torch.float64):
    """Read and check one order-k dump; returns a dict with host ``Q`` (csr) and
    ``rhs`` and device ``A`` (Q), ``rhs_device``, and for k = 2 ``K`` and ``mass``."""
    root = Path(directory)
    Q = sio.mmread(root / f"k{k}_Q.mtx").tocsr()
    rhs = np.asarray(sio.mmread(root / f"k{k}_rhs.mtx"))
    if rhs.ndim == 1:
        rhs = rhs[:, None]
    if rhs.shape[0] != Q.shape[0]:
        raise ValueError("right-hand sides do not match the system size")
    dev = torch.device("cuda", 0) if device is None else torch.device(device)
    out = dict(Q=Q, rhs=rhs, A=upload(Q, dev, dtype),
               rhs_device=torch.tensor(np.ascontiguousarray(rhs), dtype=dtype, device=dev))
    if k == 2:
        n = Q.shape[0]
        H = sio.mmread(root / "k1_Q.mtx").tocsr()
        mixed = sio.mmread(root / "k4_Q.mtx").tocsr()
        mass = mixed.diagonal()[:n].copy()
        K = (-mixed[:n, n:]).tocsr()
        if abs(H - (ss.diags(mass) + K)).max() >= 1e-10:
            raise ValueError("k1_Q is not M + K for the mass and K of k4_Q")
        expected = ss.diags(mass) + K @ ss.diags(1 / mass) @ K
        if abs(Q - expected).max() > 1e-12 * abs(Q).max():
            raise ValueError("k2_Q is not M + K M^-1 K")
        out.update(K=upload(K, dev, dtype), mass=torch.tensor(mass, dtype=dtype, device=dev),
                   K_host=K, mass_host=mass)
    return out
