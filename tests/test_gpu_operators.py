"""GPU parity for SparseOperator, the factored squared-Laplacian operator, block
Jacobi / ILU(0) and the lifted biharmonic system (reference tests
test_sparse_operator.py, test_squared_laplacian.py, test_indefinite.py,
test_biharmonic.py), against scipy, dense algebra and the CPU oracle."""

import numpy as np
import pytest
import scipy.sparse as ss
import scipy.sparse.linalg as sla
import torch

from gpu_helpers import dense, dev, host, oracle_csr, upload, vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wp():
    import warp_preconditioners as w

    return w


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32])
def test_sparse_operator_rows_aliases_and_graph(wp, dtype, lanes):
    rng = np.random.default_rng(91)
    a = ss.random(67, 67, density=0.12, random_state=rng, format="csr")
    a.data = rng.normal(size=a.nnz)
    a = a.tolil()
    a[0, :] = 0
    a = a.tocsr()
    A = wp.SparseOperator(upload(a, dtype), row_lanes=lanes)
    rtol, atol = (2e-5, 2e-6) if dtype == torch.float32 else (1e-12, 1e-12)
    for alias in ["none", "xz", "yz", "xyz"]:
        x = vec(rng.normal(size=67), dtype)
        y = x if alias == "xyz" else vec(rng.normal(size=67), dtype)
        z = x if alias in ["xz", "xyz"] else y if alias == "yz" else torch.empty_like(x)
        ref = 0.7 * (a @ host(x).astype(np.float64)) - 0.3 * host(y).astype(np.float64)
        A.matvec(x, y, z, 0.7, -0.3)
        np.testing.assert_allclose(host(z), ref, rtol=rtol, atol=atol)
    x = vec(rng.normal(size=67), dtype)
    y = torch.full_like(x, float("nan"))
    z = torch.empty_like(x)
    A.matvec(x, y, z, 1.0, 0.0)
    np.testing.assert_allclose(host(z), a @ host(x).astype(np.float64), rtol=rtol, atol=atol)
    x.fill_(float("nan"))
    y.fill_(2.0)
    A.matvec(x, y, z, 0.0, 0.5)
    np.testing.assert_array_equal(host(z), 1.0)
    x.fill_(1.0)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        A.matvec(x, y, z, 1.0, 0.0)
    for _ in range(3):
        z.fill_(-1.0)
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(z), a @ np.ones(67), rtol=rtol, atol=atol)


@pytest.mark.parametrize("lanes", [4, 8])
def test_cooperative_fsai_blocks_and_spd_solve(wp, lanes):
    from warp_preconditioners import linear

    n = 69
    a = ss.diags([-np.ones(n - 1), np.full(n, 2.1), -np.ones(n - 1)], [-1, 0, 1]).tocsr()
    raw = upload(a, block=3)
    operator = wp.SparseOperator(raw, row_lanes=8)
    pre = wp.FSAI(raw, max_row_size=8, apply_lanes=lanes)
    x = torch.zeros((23, 3), dtype=torch.float64, device=dev())
    b = torch.ones_like(x)
    linear.cg(operator, b, x, M=pre, tol=1e-12, maxiter=1000)
    np.testing.assert_allclose(a @ host(x).ravel(), 1.0, atol=1e-10)
    reference = wp.FSAI(raw, max_row_size=8)
    rng = np.random.default_rng(7)
    for alias in ["none", "xz", "yz", "xyz"]:
        u = torch.tensor(rng.normal(size=(23, 3)), device=dev())
        v = u if alias == "xyz" else torch.tensor(rng.normal(size=(23, 3)), device=dev())
        out = u if alias in ["xz", "xyz"] else v if alias == "yz" else torch.empty_like(u)
        expected = torch.empty_like(u)
        reference.matvec(u, v, expected, 0.7, -0.2)
        pre.matvec(u, v, out, 0.7, -0.2)
        np.testing.assert_allclose(host(out), host(expected), rtol=1e-12, atol=1e-12)


def test_invalid_lanes(wp):
    raw = upload(ss.eye(4))
    with pytest.raises(ValueError, match="row_lanes"):
        wp.SparseOperator(raw, row_lanes=3)
    with pytest.raises(ValueError, match="apply_lanes"):
        wp.FSAI(raw, apply_lanes=3)


def _laplacian_1d(n):
    return ss.diags([-np.ones(n - 1), np.r_[1.0, np.full(n - 2, 2.0), 1.0], -np.ones(n - 1)],
                    [-1, 0, 1]).tocsr()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8])
@pytest.mark.parametrize("mass_weight", [0.0, 0.15])
def test_factored_dirichlet_operator_and_rhs(wp, oracle, dtype, lanes, mass_weight):
    rng = np.random.default_rng(27)
    n = 67
    lap = _laplacian_1d(n)
    npd = np.float32 if dtype == torch.float32 else np.float64
    mass = rng.uniform(0.4, 2, size=n).astype(npd)
    free = rng.permutation(np.arange(3, n - 4)).astype(np.int32)
    A = wp.SquaredLaplacianOperator(upload(lap, dtype), vec(mass, dtype),
                                    torch.tensor(free, device=dev()), row_lanes=lanes,
                                    mass_weight=mass_weight)
    q = lap @ ss.diags(1 / mass.astype(np.float64)) @ lap + mass_weight * ss.diags(mass)
    reduced = q[free][:, free]
    rtol, atol = (3e-5, 3e-6) if dtype == torch.float32 else (1e-12, 1e-12)
    op = oracle.sqlap_setup(oracle_csr(oracle, lap, npd), mass, free)
    for alias in ["none", "xz", "yz", "xyz"]:
        x = vec(rng.normal(size=len(free)), dtype)
        y = x if alias == "xyz" else vec(rng.normal(size=len(free)), dtype)
        expected = 0.7 * (reduced @ host(x)) - 0.2 * host(y)
        orc = oracle.sqlap_apply(op, host(x), host(y), 0.7, -0.2, mass_weight)
        z = x if alias in ["xz", "xyz"] else y if alias == "yz" else torch.empty_like(x)
        A.matvec(x, y, z, 0.7, -0.2)
        np.testing.assert_allclose(host(z), expected, rtol=rtol, atol=atol)
        np.testing.assert_allclose(host(z), orc, rtol=rtol, atol=atol)
    bc = rng.normal(size=n).astype(npd)
    prescribed = bc.copy()
    prescribed[free] = 0
    result = A.rhs(vec(bc, dtype))
    np.testing.assert_allclose(host(result), -(q @ prescribed)[free], rtol=rtol, atol=atol)
    x = torch.ones(len(free), dtype=dtype, device=dev())
    y = torch.full_like(x, float("nan"))
    z = torch.empty_like(x)
    A.matvec(x, y, z, 1.0, 0.0)
    np.testing.assert_allclose(host(z), reduced @ np.ones(len(free)), rtol=rtol, atol=atol)
    bcw = vec(bc, dtype)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        A.matvec(x, y, z, 1.0, 0.0)
        A.rhs(bcw, out=result)
    for _ in range(2):
        z.fill_(0.0)
        result.fill_(0.0)
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(z), reduced @ np.ones(len(free)), rtol=rtol, atol=atol)
        np.testing.assert_allclose(host(result), -(q @ prescribed)[free], rtol=rtol, atol=atol)


def test_squared_validation(wp):
    L = upload(ss.eye(5))
    mass = torch.ones(5, dtype=torch.float64, device=dev())
    for weight in [-1, float("nan"), float("inf")]:
        with pytest.raises(ValueError, match="mass_weight"):
            wp.SquaredLaplacianOperator(L, mass, torch.tensor([1, 2], dtype=torch.int32,
                                                              device=dev()), mass_weight=weight)
    for free in [[0, 0], [0, 5], [-1, 2]]:
        with pytest.raises(ValueError, match="free_indices"):
            wp.SquaredLaplacianOperator(L, mass, torch.tensor(free, dtype=torch.int32, device=dev()))
    mass.zero_()
    with pytest.raises(ValueError, match="mass entries"):
        wp.SquaredLaplacianOperator(L, mass, torch.tensor([1, 2], dtype=torch.int32, device=dev()))


def test_unconstrained_smoothing(wp):
    from warp_preconditioners import linear

    n = 23
    lap = _laplacian_1d(n)
    mass = np.linspace(0.5, 2.0, n)
    target = np.sin(np.arange(n))
    weight = 0.03
    expected_matrix = lap @ ss.diags(1 / mass) @ lap + ss.diags(weight * mass)
    operator = wp.SquaredLaplacianOperator(upload(lap), vec(mass),
                                           torch.arange(n, dtype=torch.int32, device=dev()),
                                           row_lanes=4, mass_weight=weight)
    rhs = vec(weight * mass * target)
    x = vec(target)
    linear.cg(operator, rhs, x, tol=1e-12, maxiter=300)
    np.testing.assert_allclose(host(x), sla.spsolve(expected_matrix, weight * mass * target),
                               rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("size", [1, 2, 3, 4])
def test_block_factors_and_aliases(wp, dtype, size):
    rng = np.random.default_rng(61)
    n = 5
    blocks = rng.normal(size=(n, n, size, size)) * 0.03
    for i in range(n):
        blocks[i, i] = np.diag(np.arange(1, size + 1) * np.where(np.arange(size) % 2, -1.0, 1.0))
    a = blocks.transpose(0, 2, 1, 3).reshape(n * size, n * size)
    A = upload(a, dtype, size)
    pre = wp.BlockILU0(A, factor_sweeps=2 * n + 2, solve_sweeps=n)
    jac = wp.BlockJacobi(A)
    expected = np.linalg.inv(a)
    tol = dict(rtol=2e-5, atol=1e-6) if dtype == torch.float32 else dict(rtol=1e-12, atol=1e-12)
    for alias in ["none", "xz", "yz", "xyz"]:
        x = torch.tensor(rng.normal(size=(n, size)), dtype=dtype, device=dev())
        y = x if alias == "xyz" else torch.tensor(rng.normal(size=(n, size)), dtype=dtype, device=dev())
        z = x if alias in ["xz", "xyz"] else y if alias == "yz" else torch.empty_like(x)
        target = 0.7 * (expected @ host(x).ravel()) - 0.2 * host(y).ravel()
        pre.matvec(x, y, z, 0.7, -0.2)
        np.testing.assert_allclose(host(z).ravel(), target, **tol)
    x = torch.tensor(rng.normal(size=(n, size)), dtype=dtype, device=dev())
    y = torch.empty_like(x)
    jac.matvec(x, y, y, 1.0, 0.0)
    target = np.stack([np.linalg.solve(blocks[i, i], host(x)[i]) for i in range(n)])
    np.testing.assert_allclose(host(y), target, rtol=1e-6, atol=1e-6)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        pre.matvec(x, y, y, 1.0, 0.0)
    for _ in range(2):
        y.fill_(0.0)
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(y).ravel(), expected @ host(x).ravel(), **tol)


def test_bad_block_pivot(wp):
    with pytest.raises(ValueError, match="pivot"):
        wp.BlockJacobi(upload(np.diag([1.0, 0.0]), block=2))


def test_sparse_ilu_pattern_and_linearity(wp, oracle):
    rng = np.random.default_rng(83)
    n, size = 9, 2
    blocks = {}
    for i in range(n):
        for j in range(n):
            if i == j:
                blocks[i, j] = np.diag([3.0, -2.0])
            elif abs(i - j) == 1 or (i + j) % 5 == 0:
                blocks[i, j] = rng.normal(size=(size, size)) * 0.08
    a = np.zeros((size * n, size * n))
    for (i, j), value in blocks.items():
        a[size * i:size * (i + 1), size * j:size * (j + 1)] = value
    A = upload(a, block=size)
    pre = wp.BlockILU0(A, factor_sweeps=2 * n + 2, solve_sweeps=3)
    factors, inv = oracle.block_ilu0(oracle_csr(oracle, a, block=size), factor_sweeps=2 * n + 2)
    np.testing.assert_allclose(host(pre.factors), factors, rtol=1e-12, atol=1e-12)

    def apply(v):
        x = torch.tensor(v, device=dev())
        y = torch.empty_like(x)
        pre.matvec(x, y, y, 1.0, 0.0)
        return host(y)

    x, y = rng.normal(size=(2, n, size))
    np.testing.assert_allclose(apply(0.3 * x + 0.7 * y), 0.3 * apply(x) + 0.7 * apply(y),
                               rtol=1e-12, atol=1e-12)
    ref = oracle.block_ilu0_apply(oracle_csr(oracle, a, block=size), factors, inv, x, solve_sweeps=3)
    np.testing.assert_allclose(apply(x), ref, rtol=1e-12, atol=1e-12)
    xw = torch.full((n, size), float("nan"), dtype=torch.float64, device=dev())
    yw = torch.tensor(y, device=dev())
    pre.matvec(xw, yw, xw, 0.0, 0.7)
    np.testing.assert_allclose(host(xw), 0.7 * y)


@pytest.mark.parametrize("tau", [0.02, 2.0])
def test_lifted_biharmonic(wp, tau):
    rng = np.random.default_rng(13)
    k = rng.normal(size=(10, 10))
    k = k @ k.T
    mass = rng.uniform(0.4, 2.0, size=10)
    system = wp.BiharmonicSystem(upload(k), vec(mass), tau)
    pre = system.preconditioner(max_row_size=10, kap_tolerance=0)
    h_inv = np.linalg.inv(np.diag(mass) + np.sqrt(tau) * k)
    v = torch.tensor(rng.normal(size=(10, 2)), device=dev())
    out = torch.empty_like(v)
    pre.matvec(v, out, out, 1.0, 0.0)
    np.testing.assert_allclose(host(out), h_inv @ host(v), atol=1e-11)
    actual = dense(system.matrix)
    expect = np.zeros((20, 20))
    expect[::2, ::2] = np.diag(mass)
    expect[1::2, 1::2] = np.diag(mass)
    expect[::2, 1::2] = -np.sqrt(tau) * k
    expect[1::2, ::2] = np.sqrt(tau) * k
    np.testing.assert_allclose(actual, expect)
    b = system.rhs(vec(rng.normal(size=10)))
    assert tuple(b.shape) == (10, 2)
    np.testing.assert_array_equal(host(system.solution(b)), host(b)[:, 0])


@pytest.mark.parametrize("tau", [0.02, 2.0])
def test_lifted_biharmonic_gmres(wp, tau):
    # tests/test_biharmonic.py:13-43: GMRES on the lifted system recovers the Q solve.
    from warp_preconditioners import linear

    rng = np.random.default_rng(13)
    k = rng.normal(size=(10, 10))
    k = k @ k.T
    mass = rng.uniform(0.4, 2.0, size=10)
    b_np = rng.normal(size=10)
    system = wp.BiharmonicSystem(upload(k), vec(mass), tau)
    Q = np.diag(mass) + tau * (k / mass) @ k
    b = system.rhs(vec(b_np))
    x = torch.zeros_like(b)
    pre = system.preconditioner(max_row_size=10, kap_tolerance=0)
    linear.gmres(system.matrix, b, x, M=pre, tol=1e-12, maxiter=100, restart=20)
    u = host(system.solution(x))
    np.testing.assert_allclose(u, np.linalg.solve(Q, b_np), rtol=1e-9, atol=1e-10)
    np.testing.assert_allclose(Q @ u, b_np, rtol=1e-9, atol=1e-9)


def test_block_ilu_gmres_on_indefinite_blocks(wp):
    # tests/test_indefinite.py:127-139 pattern: GMRES + BlockILU0 on an indefinite
    # 2x2-block system with a manufactured solution.
    from warp_preconditioners import linear

    rng = np.random.default_rng(7)
    n = 12
    blocks = rng.normal(size=(n, n, 2, 2)) * 0.05
    for i in range(n):
        blocks[i, i] = np.array([[2.0, 0.3], [0.3, -1.5]])
    a = blocks.transpose(0, 2, 1, 3).reshape(2 * n, 2 * n)
    truth = rng.normal(size=(n, 2))
    A = upload(a, block=2)
    b = torch.tensor((a @ truth.ravel()).reshape(n, 2), device=dev())
    x = torch.zeros_like(b)
    linear.gmres(A, b, x, M=wp.BlockILU0(A, factor_sweeps=20, solve_sweeps=n), tol=1e-11,
                 maxiter=60, restart=15)
    np.testing.assert_allclose(host(x), truth, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("sweeps", [0, 1, 2, 3])
@pytest.mark.parametrize("relaxation", [1.0, 0.7])
def test_block_ilu_few_sweeps_match_oracle(wp, oracle, sweeps, relaxation):
    # Away from the fixed point (0-3 sweeps, as configuration 4 runs 3; relaxation
    # 1 and 0.7): the device factors, U-diagonal inverses and apply equal the
    # oracle's (itself pinned by a numpy restatement of one synchronous sweep).
    rng = np.random.default_rng(90 + sweeps)
    n, size = 60, 3
    blocks = {}
    for i in range(n):
        for j in range(n):
            if i == j:
                blocks[i, j] = rng.normal(size=(size, size)) * 0.2 + np.diag([2.0, 2.5, 3.0])
            elif abs(i - j) <= 2 or (i * 7 + j * 3) % 11 == 0:
                blocks[i, j] = rng.normal(size=(size, size)) * 0.35
    a = np.zeros((size * n, size * n))
    for (i, j), value in blocks.items():
        a[size * i:size * (i + 1), size * j:size * (j + 1)] = value
    A = upload(a, block=size)
    pre = wp.BlockILU0(A, factor_sweeps=sweeps, solve_sweeps=4, relaxation=relaxation)
    Ao = oracle_csr(oracle, a, block=size)
    factors, inv = oracle.block_ilu0(Ao, factor_sweeps=sweeps, relaxation=relaxation)
    np.testing.assert_allclose(host(pre.factors), factors, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(host(pre.inverse_diagonal), inv, rtol=1e-12, atol=1e-13)
    other, _ = oracle.block_ilu0(Ao, factor_sweeps=sweeps + 1, relaxation=relaxation)
    assert np.abs(other - factors).max() > 1e-6  # one sweep more would not pass
    x = rng.normal(size=(n, size))
    y = rng.normal(size=(n, size))
    z = torch.tensor(y, device=dev())
    pre.matvec(torch.tensor(x, device=dev()), z, z, 0.9, -0.4)
    ref = oracle.block_ilu0_apply(Ao, factors, inv, x, y, 0.9, -0.4, solve_sweeps=4)
    np.testing.assert_allclose(host(z), ref, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("sweeps,solve_sweeps", [(3, 4), (20, 8)])
def test_block_ilu_on_reduced_beam_matches_oracle(wp, oracle, sweeps, solve_sweeps):
    # The reduced elasticity beam of tests/test_oracle_ilu_beam.py (where the oracle
    # shows CG diverging with 3 sweeps and converging with 20/8): the device factors
    # and apply equal the oracle's, so the port computes the same preconditioner.
    from test_oracle_ilu_beam import beam
    from warp_preconditioners.sparse import BsrMatrix

    K, b = beam()
    nb = K.shape[0] // 3
    A = BsrMatrix(nb, nb, (3, 3), torch.tensor(K.indptr.astype(np.int32), device=dev()),
                  torch.tensor(K.indices.astype(np.int32), device=dev()),
                  torch.tensor(np.ascontiguousarray(K.data), device=dev()))
    pre = wp.BlockILU0(A, factor_sweeps=sweeps, solve_sweeps=solve_sweeps)
    Ao = oracle.CSR(K.indptr.astype(np.int32), K.indices.astype(np.int32),
                    np.ascontiguousarray(K.data), nb, nb, 3)
    factors, inv = oracle.block_ilu0(Ao, factor_sweeps=sweeps)
    scale = np.abs(factors).max()
    np.testing.assert_allclose(host(pre.factors), factors, rtol=1e-11, atol=1e-13 * scale)
    z = torch.empty(nb * 3, dtype=torch.float64, device=dev())
    pre.matvec(torch.tensor(b, device=dev()), z, z, 1.0, 0.0)
    ref = oracle.block_ilu0_apply(Ao, factors, inv, b, solve_sweeps=solve_sweeps).ravel()
    np.testing.assert_allclose(host(z), ref, rtol=1e-10, atol=1e-13 * np.abs(ref).max())
