"""GPU CG/CR: iteration counts against the reference's recorded grid results and the
oracle, batched right-hand sides against single solves."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host, oracle_csr, upload, vec

pytestmark = pytest.mark.gpu


def grid_problem(size):
    # benchmarks/synthetic.py:24-30 (Q = h^2 I + K^2/h^2, seed 13).
    d = ss.diags([-np.ones(size - 1), 2 * np.ones(size), -np.ones(size - 1)], [-1, 0, 1])
    k = ss.kronsum(d, d).tocsr()
    mass = 1 / (size + 1) ** 2
    q = (mass * ss.eye(size * size) + k @ k / mass).tocsr()
    exact = np.random.default_rng(13).normal(size=size * size)
    return q, q @ exact


@pytest.mark.parametrize("width,count", [(1, 600), (4, 300), (8, 175), (16, 125)])
def test_grid32_iteration_counts_match_recorded(width, count):
    # results/grid32.json: float64, tol 1e-11, check_every 25.
    from warp_preconditioners import FSAI, linear

    q, b = grid_problem(32)
    Q = upload(q)
    M = linear.preconditioner(Q, "diag") if width == 1 else FSAI(Q, max_row_size=width)
    x = torch.zeros(len(b), dtype=torch.float64, device=dev())
    it, res, atol = linear.cg(Q, vec(b), x, M=M, tol=1e-11, maxiter=20000, check_every=25)
    assert it == count
    xs = host(x)
    assert np.linalg.norm(b - q @ xs) / np.linalg.norm(b) < 2e-11


@pytest.mark.parametrize("method", ["cg", "cr"])
def test_krylov_matches_oracle_iterations(oracle, method):
    from warp_preconditioners import FSAI, linear

    q, b = grid_problem(24)
    Q = upload(q)
    P = FSAI(Q, max_row_size=8)
    x = torch.zeros(len(b), dtype=torch.float64, device=dev())
    solve = linear.cg if method == "cg" else linear.cr
    it, res, _ = solve(Q, vec(b), x, M=P, tol=1e-10, maxiter=5000, check_every=1)
    Po = oracle.fsai(oracle_csr(oracle, q), max_row_size=8)
    xo, ito, reso = oracle.krylov(method, b, A=oracle_csr(oracle, q), P=Po, tol=1e-10,
                                  maxiter=5000, check_every=1)
    assert abs(it - ito) <= max(2, 0.02 * ito)
    np.testing.assert_allclose(host(x), xo, rtol=1e-8, atol=1e-10 * np.abs(xo).max())


def test_exact_count_device_loop():
    from warp_preconditioners import FSAI, linear

    q, b = grid_problem(16)
    Q = upload(q)
    P = FSAI(Q, max_row_size=8)
    x = torch.zeros(len(b), dtype=torch.float64, device=dev())
    nit, rsq, atol = linear.cg(Q, vec(b), x, M=P, tol=1e-10, maxiter=2000, check_every=0)
    x2 = torch.zeros_like(x)
    it1, _, _ = linear.cg(Q, vec(b), x2, M=P, tol=1e-10, maxiter=2000, check_every=1)
    assert int(nit.cpu()[0]) == it1
    np.testing.assert_array_equal(host(x), host(x2))
    assert float(rsq.cpu()[0]) ** 0.5 <= atol


@pytest.mark.parametrize("prec", ["jacobi", "fsai-bf16"])
def test_refined_solve_reaches_true_residual(prec):
    # K (x) C with eigenvalue contrast 1e6 (eps * cond ~ 1e-7): float32 inner CG with
    # float64 refinement reaches a compensated true relative residual <= 1e-8 and the
    # direct solution.
    import scipy.sparse as ss
    import scipy.sparse.linalg as sla

    import warp_preconditioners as wp
    from warp_preconditioners import linear, workloads
    from warp_preconditioners.sparse import BsrMatrix

    off, col, val, _, k = workloads.kron_coupling(12, 1.0e6, seed=3)
    nb = len(off) - 1

    def bsr(dt):
        return BsrMatrix(nb, nb, (3, 3), torch.tensor(off, device=dev()),
                         torch.tensor(col, device=dev()), torch.tensor(val, dtype=dt, device=dev()))

    A64, A32 = bsr(torch.float64), bsr(torch.float32)
    op64, op32 = wp.SparseOperator(A64, row_lanes=1), wp.SparseOperator(A32, row_lanes=1)
    # pivot tolerance as the reference passes for this system (block_challenge.py:101)
    P = wp.BlockJacobi(A32, pivot_tolerance=1e-15) if prec == "jacobi" else wp.FSAI(A64, max_row_size=16,
                                                              factor_dtype=torch.bfloat16)
    rng = np.random.default_rng(9)
    bh = rng.standard_normal(nb * 3)
    b = torch.tensor(bh, device=dev())
    x = torch.zeros_like(b)
    out = linear.solve_refined("cg", op64, b, x, M=P, inner=op32, tol=1e-8, inner_tol=1e-4)
    assert out["converged"] and out["relative_residual"] <= 1e-8 and out["outer"] >= 2
    dense = ss.bsr_matrix((val, col, off), shape=(nb * 3, nb * 3)).tocsc()
    xs = sla.spsolve(dense, bh)
    assert np.linalg.norm(host(x) - xs) / np.linalg.norm(xs) < 1e-5
