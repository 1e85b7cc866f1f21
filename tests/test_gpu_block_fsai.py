"""Block FSAI on the device: bit-identical to the oracle, and the reference's own
dense checks (tests/test_fsai_candidates.py:20-55, test_block_challenge.py:65-90)."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host

pytestmark = pytest.mark.gpu


def _spd(rng, n, scaled=True):
    r = rng.normal(size=(n, n))
    a = r @ r.T + np.eye(n)
    if scaled:
        s = np.geomspace(0.01, 100, n)
        a *= s[:, None] * s[None, :]
    return a


def _upload(a, bs, dtype=torch.float64):
    from warp_preconditioners.sparse import from_scipy

    return from_scipy(ss.csr_matrix(a), dev(), dtype, (bs, bs))


def _rows_check(g, a, bs, nblocks, rtol=2e-11, atol=1e-10):
    for i in range(nblocks):
        row = g[bs * i: bs * (i + 1)]
        support = np.flatnonzero(np.any(row != 0, axis=0))
        rhs = np.zeros((len(support), bs))
        rhs[-bs:, :] = np.eye(bs)
        z = np.linalg.solve(a[np.ix_(support, support)], rhs)
        expected = np.linalg.solve(np.linalg.cholesky(z[-bs:, :]), z.T)
        np.testing.assert_allclose(row[:, support], expected, rtol=rtol, atol=atol)
        np.testing.assert_allclose(row @ a @ row.T, np.eye(bs), atol=2e-12)


@pytest.mark.parametrize("bs", [2, 3])
@pytest.mark.parametrize("blocks", [1, 2, 4])
@pytest.mark.parametrize("cls", ["BlockFSAI", "BsrBlockFSAI"])
def test_block_principal_inverse(bs, blocks, cls):
    import warp_preconditioners as wp

    rng = np.random.default_rng(712)
    n = 4 * bs
    a = _spd(rng, n)
    pre = getattr(wp, cls)(_upload(a, bs), max_blocks=blocks, factor_dtype=torch.float64)
    g = pre.G.to_dense()
    assert np.linalg.eigvalsh(g.T @ g).min() > 0
    _rows_check(g, a, bs, 4)
    if blocks == 4:
        np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=2e-11, atol=1e-10)
    x = torch.tensor(rng.normal(size=n), device=dev())
    expected = 0.7 * g.T @ g @ host(x) - 0.2 * host(x)
    pre.matvec(x, x, x, 0.7, -0.2)
    np.testing.assert_allclose(host(x), expected, rtol=2e-11, atol=1e-10)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_block_factor_bitwise_equals_oracle(oracle, dtype):
    import warp_preconditioners as wp

    rng = np.random.default_rng(3)
    bs, nb = 3, 60
    n = bs * nb
    mask = rng.random((nb, nb)) < 0.15
    mask = mask | mask.T
    np.fill_diagonal(mask, True)
    a = _spd(rng, n, scaled=False) * np.kron(mask, np.ones((bs, bs))) + n * np.eye(n)
    npd = np.float64 if dtype == torch.float64 else np.float32
    P = oracle.block_fsai(oracle.CSR.from_scipy(a, npd, bs), max_blocks=4, factor_dtype=npd)
    pre = wp.BlockFSAI(_upload(a, bs, dtype), max_blocks=4, factor_dtype=dtype)
    np.testing.assert_array_equal(host(pre.G.offsets), P["G"].off)
    np.testing.assert_array_equal(host(pre.G.columns), P["G"].col)
    np.testing.assert_array_equal(host(pre.G.values), P["G"].val)


def test_block_refit_selected_principal_inverse(oracle):
    import warp_preconditioners as wp

    rng = np.random.default_rng(992)
    old = _spd(rng, 12, scaled=False) + 2 * np.eye(12)
    new = _spd(rng, 12, scaled=False) + 4 * np.eye(12)
    for cls in ("BlockFSAI", "BsrBlockFSAI"):
        pre = getattr(wp, cls)(_upload(old, 3), max_blocks=2, factor_dtype=torch.float64,
                               reuse_pattern=True)
        columns = host(pre.G.columns).copy()
        pre.update(_upload(new, 3))
        np.testing.assert_array_equal(columns, host(pre.G.columns))
        g = pre.G.to_dense()
        _rows_check(g, new, 3, 4, rtol=2e-12, atol=2e-13)
        x = torch.tensor(rng.normal(size=12), device=dev())
        expected = g.T @ g @ host(x)
        pre.matvec(x, x, x, 1.0, 0.0)
        np.testing.assert_allclose(host(x), expected, rtol=2e-12, atol=2e-13)
    # the device refit equals the oracle's (pattern read back from the sorted rows)
    P = oracle.block_fsai(oracle.CSR.from_scipy(old, np.float64, 3), max_blocks=2)
    Q = oracle.block_fsai(oracle.CSR.from_scipy(new, np.float64, 3), max_blocks=2, refit_from=P)
    pre = wp.BlockFSAI(_upload(old, 3), max_blocks=2, factor_dtype=torch.float64,
                       reuse_pattern=True).update(_upload(new, 3))
    np.testing.assert_array_equal(host(pre.G.values), Q["G"].val)


def test_block_fsai_cg_on_block_coupling():
    # K (x) C: block FSAI drives the fused CG path to the direct solution.
    import scipy.sparse.linalg as sla

    import warp_preconditioners as wp
    from warp_preconditioners import linear, workloads
    from warp_preconditioners.sparse import BsrMatrix

    off, col, val, _, _ = workloads.kron_coupling(10, 1.0e3, seed=1)
    nb = len(off) - 1
    A = BsrMatrix(nb, nb, (3, 3), torch.tensor(off, device=dev()), torch.tensor(col, device=dev()),
                  torch.tensor(val, device=dev()))
    pre = wp.BlockFSAI(A, max_blocks=4, factor_dtype=torch.float32)
    b = torch.tensor(np.random.default_rng(2).standard_normal(nb * 3), device=dev())
    x = torch.zeros_like(b)
    its, res, atol = linear.cg(A, b, x, M=pre, tol=1e-10, maxiter=5000)
    dense = ss.bsr_matrix((val, col, off), shape=(nb * 3, nb * 3)).tocsc()
    xs = sla.spsolve(dense, host(b))
    assert np.linalg.norm(host(x) - xs) / np.linalg.norm(xs) < 1e-7
    its_j, _, _ = linear.cg(A, b, torch.zeros_like(b), M=wp.BlockJacobi(A), tol=1e-10,
                            maxiter=5000)
    assert its <= its_j
