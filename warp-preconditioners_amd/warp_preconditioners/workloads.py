"""Seeded synthetic inputs for the benchmark configurations (host numpy/scipy).

* ``laplacian_2d(64)`` -- five-point Dirichlet Laplacian on a 64x64 interior grid
  (diagonal 4, off-diagonal -1, row-major i*64+j), 4,096 unknowns.  The reference
  builds the same operator in benchmarks/fsai_dtype_regression.py:57-75.
* ``kuhn_grid(n, jitter, seed)`` -- n^3 vertices on the unit cube, every cube split
  into 6 tetrahedra along the coordinate paths of the permutations of (x, y, z)
  (the Freudenthal/Kuhn split of benchmarks/block_refinement.py:21-34), interior
  vertices jittered uniformly by +-jitter*h.  Vertex id = (x*n + y)*n + z.
* ``p1_laplacian(vertices, tets)`` -- P1 stiffness K_ij = sum_T vol*grad(phi_i).grad(phi_j)
  (the scalar term of benchmarks/elasticity_problem.py:33-61) and lumped mass vol/4.
* ``kron_coupling(n, contrast, seed)`` -- 7-point Dirichlet Laplacian on n^3 (x fastest,
  id x + n*y + n^2*z as benchmarks/block_challenge.py:29-58) Kronecker a 3x3 SPD C with
  eigenvalues (1, 1, contrast) rotated by a seeded Haar rotation, as 3x3 BSR.
* ``p1_elasticity(vertices, tets, fixed, young, poisson)`` -- 3x3-block isotropic
  linear elasticity (benchmarks/elasticity_problem.py:15-158) with Dirichlet
  elimination of `fixed` vertices and a lumped body force.
"""

from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as ss


def laplacian_2d(side: int = 64) -> ss.csr_matrix:
    ids = np.arange(side * side).reshape(side, side)
    horizontal = np.column_stack((ids[:, :-1].ravel(), ids[:, 1:].ravel()))
    vertical = np.column_stack((ids[:-1].ravel(), ids[1:].ravel()))
    edges = np.concatenate((horizontal, vertical))
    n = side * side
    rows = np.concatenate((np.arange(n), edges[:, 0], edges[:, 1]))
    cols = np.concatenate((np.arange(n), edges[:, 1], edges[:, 0]))
    vals = np.concatenate((np.full(n, 4.0), np.full(2 * len(edges), -1.0)))
    a = ss.csr_matrix((vals, (rows, cols)), shape=(n, n))
    a.sum_duplicates()
    a.sort_indices()
    return a


def kuhn_grid(n, jitter=0.15, seed=0, shape=None):
    """Vertices, tetrahedra (int32 [T, 4]) and the boundary mask of a jittered grid."""
    shape = (n, n, n) if shape is None else tuple(shape)
    h = 1.0 / (min(shape) - 1) if shape[0] == shape[1] == shape[2] else 1.0 / (shape[1] - 1)
    idx = np.indices(shape).reshape(3, -1).T
    vertices = idx.astype(np.float64) * h
    boundary = np.any((idx == 0) | (idx == np.array(shape) - 1), axis=1)
    if jitter:
        rng = np.random.default_rng(seed)
        noise = rng.uniform(-jitter * h, jitter * h, size=vertices.shape)
        vertices[~boundary] += noise[~boundary]
    base = np.indices(tuple(s - 1 for s in shape)).reshape(3, -1).T
    cells = []
    for order in itertools.permutations(range(3)):
        corners = [np.zeros(3, dtype=np.int64)]
        for axis in order:
            c = corners[-1].copy()
            c[axis] += 1
            corners.append(c)
        cidx = base[:, None, :] + np.array(corners)[None, :, :]
        cells.append(np.ravel_multi_index(cidx.transpose(2, 0, 1), shape))
    tets = np.concatenate(cells).astype(np.int32)
    return vertices, tets, boundary


def _gradients(vertices, tets):
    v = vertices[tets]  # [T, 4, 3]
    dm = (v[:, 1:, :] - v[:, :1, :]).transpose(0, 2, 1)  # columns are edges
    det = np.linalg.det(dm)
    inv = np.linalg.inv(dm)  # rows are gradients of phi_1..3
    grads = np.concatenate([-inv.sum(axis=1, keepdims=True), inv], axis=1)  # [T, 4, 3]
    return np.abs(det) / 6.0, grads


def p1_laplacian(vertices, tets):
    """P1 stiffness K (csr, float64) and lumped mass (vol/4 per vertex)."""
    n = len(vertices)
    vol, g = _gradients(vertices, tets)
    local = vol[:, None, None] * np.einsum("tia,tja->tij", g, g)
    rows = np.repeat(tets, 4, axis=1).reshape(-1)
    cols = np.tile(tets, (1, 4)).reshape(-1)
    K = ss.csr_matrix((local.reshape(-1), (rows, cols)), shape=(n, n))
    K.sum_duplicates()
    K.sort_indices()
    mass = np.bincount(tets.reshape(-1), weights=np.repeat(vol / 4.0, 4), minlength=n)
    return K, mass


def squared_operator(K, mass, alpha=1.0):
    """Assembled Q = alpha*M + K M^-1 K (host SpGEMM), for FSAI setup."""
    q = (K @ ss.diags(1.0 / mass) @ K + alpha * ss.diags(mass)).tocsr()
    q.sum_duplicates()
    q.sort_indices()
    return q


def biharmonic_rhs(vertices, mass, alpha=1.0, noise=1e-2, seed=0):
    """alpha*M*f for f = x, y and 0.5x + sin(3y) + z^2 + xz, each plus N(0, noise^2)."""
    x, y, z = vertices.T
    fields = np.stack([x, y, 0.5 * x + np.sin(3 * y) + z * z + x * z], axis=1)
    fields = fields + np.random.default_rng(seed).normal(0.0, noise, size=fields.shape)
    return alpha * mass[:, None] * fields


def haar_rotation(seed):
    q, r = np.linalg.qr(np.random.default_rng(seed).normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def kron_coupling(n=48, contrast=1.0e6, seed=0):
    """(block offsets, columns, values [nnzb, 3, 3], C) of K (x) C on an n^3 grid."""
    d = ss.diags([-np.ones(n - 1), 2 * np.ones(n), -np.ones(n - 1)], [-1, 0, 1])
    eye = ss.eye(n)
    k = (ss.kron(ss.kron(eye, eye), d) + ss.kron(ss.kron(eye, d), eye)
         + ss.kron(ss.kron(d, eye), eye)).tocsr()
    k.sort_indices()
    r = haar_rotation(seed)
    c = r @ np.diag([1.0, 1.0, contrast]) @ r.T
    values = k.data[:, None, None] * c[None, :, :]
    return k.indptr.astype(np.int32), k.indices.astype(np.int32), values, c, k


def p1_elasticity(vertices, tets, fixed, young=1.0, poisson=0.45, density=1.0,
                  gravity=(0.0, 0.0, -1.0)):
    """Free-DOF 3x3-block stiffness (scipy bsr) and body-force RHS [nfree, 3]."""
    n = len(vertices)
    mu = young / (2 * (1 + poisson))
    lam = young * poisson / ((1 + poisson) * (1 - 2 * poisson))
    vol, g = _gradients(vertices, tets)
    free = np.flatnonzero(~fixed)
    index = np.full(n, -1, dtype=np.int64)
    index[free] = np.arange(len(free))
    gi = g[:, :, None, :, None]  # [T,4,1,3,1]
    gj = g[:, None, :, None, :]  # [T,1,4,1,3]
    dot = np.einsum("tia,tja->tij", g, g)
    blocks = (lam * gi * gj + mu * g[:, None, :, :, None] * g[:, :, None, None, :]
              + mu * dot[:, :, :, None, None] * np.eye(3))
    blocks *= vol[:, None, None, None, None]
    ri = index[np.repeat(tets, 4, axis=1).reshape(-1)]
    cj = index[np.tile(tets, (1, 4)).reshape(-1)]
    keep = (ri >= 0) & (cj >= 0)
    nb = len(free)
    data = blocks.reshape(-1, 3, 3)[keep]
    coo_rows = (ri[keep][:, None] * 3 + np.arange(3)[None, :])
    rows = np.repeat(coo_rows, 3, axis=1).reshape(-1)
    cols = np.tile(cj[keep][:, None] * 3 + np.arange(3)[None, :], (1, 3)).reshape(-1)
    K = ss.csr_matrix((data.reshape(-1), (rows, cols)), shape=(3 * nb, 3 * nb))
    K.sum_duplicates()
    K = ss.bsr_matrix(K, blocksize=(3, 3))
    K.sort_indices()
    mass = np.bincount(tets.reshape(-1), weights=np.repeat(vol / 4.0, 4), minlength=n)
    rhs = density * mass[free][:, None] * np.asarray(gravity)[None, :]
    return K, rhs, mass, free
