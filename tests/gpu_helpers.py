"""Shared helpers for the GPU parity tests (inputs and oracle conversions)."""

import numpy as np
import scipy.sparse as ss
import torch

NP = {torch.float16: np.float16, torch.float32: np.float32, torch.float64: np.float64}


def dev():
    return torch.device("cuda", 0)


def upload(a, dtype=torch.float64, block=1):
    from warp_preconditioners.sparse import from_scipy

    return from_scipy(ss.csr_matrix(a), dev(), dtype, (block, block))


def vec(values, dtype=torch.float64):
    return torch.tensor(np.asarray(values), dtype=dtype, device=dev())


def dense(m):
    return m.to_dense()


def host(t):
    return t.detach().cpu().numpy()


def oracle_csr(cpu, a, dtype=np.float64, block=1):
    return cpu.CSR.from_scipy(ss.csr_matrix(a), dtype, block)


def random_spd(rng, n, density, scale_range=None):
    a = ss.random(n, n, density=density, random_state=rng).toarray()
    a = a + a.T
    np.fill_diagonal(a, np.abs(a).sum(axis=1) + 0.5)
    if scale_range is not None:
        s = np.geomspace(*scale_range, n)
        a = s[:, None] * a * s[None, :]
    return a


def laplacian_2d(side):
    """Five-point Dirichlet Laplacian on a side x side interior grid (diag 4, -1)."""
    d = ss.diags([-np.ones(side - 1), 2 * np.ones(side), -np.ones(side - 1)], [-1, 0, 1])
    return ss.kronsum(d, d).tocsr()
