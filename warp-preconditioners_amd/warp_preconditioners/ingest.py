"""Ingest of the reference's benchmark dumps: Matrix Market systems and meshes.

``load_dump(directory, k)`` reads the files the reference's dragon benchmark
reads (benchmarks/dragon.py:79-96): ``k{k}_Q.mtx`` (the order-k system) and
``k{k}_rhs.mtx``; for k = 2 also ``k1_Q.mtx`` (H = M + K) and ``k4_Q.mtx`` (the
mixed system, whose leading diagonal is the lumped mass and whose upper-right
block is -K), checks H = M + K and Q = M + K M^-1 K as the benchmark does, and
uploads everything as device matrices.  ``load_mesh_npz(directory)`` reads the
tetrahedral mesh dumps of the reference's elasticity benchmarks (``mesh.npz`` with
``vertices``, ``tets``, ``fixed`` and optionally ``forces``, as
benchmarks/simjeb_problem.py:117-119 writes them, and ``mesh.json`` with the
material) and assembles the free-vertex P1 elasticity system as 3x3 BSR on the
device.  Host parsing is numpy's and scipy's; the data themselves are not shipped
(the benchmarks download them separately).
"""

from __future__ import annotations

import json
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


# Material defaults of the reference's elasticity benchmark when mesh.json omits a
# field (benchmarks/elasticity.py:74-77).
MESH_DEFAULTS = {"young_pa": 1.0e7, "poisson": 0.35, "density_kg_m3": 1000.0}
GRAVITY = (0.0, 0.0, -9.81)


def load_mesh_npz(directory, device=None, dtype=torch.float64):
    """Read ``mesh.npz`` (+ ``mesh.json``) and assemble the P1 elasticity system on the
    free vertices (benchmarks/elasticity.py:72-81 with elasticity_problem.py's
    assembly): stiffness as a device 3x3 BSR matrix ``A`` (scipy copy ``K_host``),
    right-hand side ``rhs`` [nfree, 3] -- the dump's nodal ``forces`` on the free
    vertices, or density x lumped volume x gravity (0, 0, -9.81) when it has none --
    and ``rhs_device`` (interleaved), ``free`` (vertex ids), ``mass`` (lumped volume
    per vertex) and the material used."""
    from . import workloads

    root = Path(directory)
    with np.load(root / "mesh.npz") as d:  # arrays only (no pickled objects)
        vertices = np.asarray(d["vertices"], dtype=np.float64)
        tets = np.asarray(d["tets"], dtype=np.int64)
        fixed = np.asarray(d["fixed"], dtype=bool)
        forces = np.asarray(d["forces"], dtype=np.float64) if "forces" in d.files else None
    if vertices.ndim != 2 or vertices.shape[1] != 3:
        raise ValueError("vertices must be an [n, 3] array")
    if tets.ndim != 2 or tets.shape[1] != 4 or tets.min() < 0 or tets.max() >= len(vertices):
        raise ValueError("tets must be an [m, 4] array of vertex indices")
    if fixed.shape != (len(vertices),):
        raise ValueError("fixed must hold one flag per vertex")
    if forces is not None and forces.shape != vertices.shape:
        raise ValueError("forces must be an [n, 3] array")
    meta_path = root / "mesh.json"
    meta = json.loads(meta_path.read_text()) if meta_path.exists() else {}
    mat = {k: float(meta.get(k, v)) for k, v in MESH_DEFAULTS.items()}
    K, rhs, mass, free = workloads.p1_elasticity(
        vertices, tets, fixed, young=mat["young_pa"], poisson=mat["poisson"],
        density=mat["density_kg_m3"], gravity=GRAVITY)
    if forces is not None:
        rhs = forces[free]
    dev = torch.device("cuda", 0) if device is None else torch.device(device)
    A = from_scipy(K, dev, dtype, (3, 3))
    return dict(A=A, K_host=K, rhs=rhs, free=free, mass=mass, material=mat,
                rhs_device=torch.tensor(np.ascontiguousarray(rhs).reshape(-1), dtype=dtype,
                                        device=dev))
