"""GPU parity of the mixed harmonic systems and their block preconditioners.

The dense checks follow tests/test_indefinite.py:58-139 and :195-229; the
larger cases compare the HIP path with the oracle (oracle.cpu.mixed_system,
matching_schur_apply, shifted_block_apply) on the same inputs.  Tolerances: the
K products run four lanes per row (a different summation order from the
oracle's serial sums), so f64 results agree to 1e-12 and f32 to 2e-5 of the
largest entry.
"""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host, laplacian_2d, upload, vec

pytestmark = pytest.mark.gpu


def _graph(n, seed=7):
    rng = np.random.default_rng(seed)
    edges = rng.uniform(0.1, 0.3, size=(n, n))
    edges = (edges + edges.T) / 2
    k = np.diag(edges.sum(axis=1)) - edges
    mass = rng.uniform(0.7, 1.2, size=n)
    return rng, k, mass


def _expected(k, mass, alpha, order):
    M = np.diag(mass)
    zero = np.zeros_like(M)
    full = (np.block([[alpha * M, -k], [-k, -M]]) if order == 2
            else np.block([[alpha * M, zero, -k], [zero, -k, -M], [-k, -M, zero]]))
    n = len(mass)
    perm = np.arange(n * order).reshape(order, n).T.ravel()
    return full[np.ix_(perm, perm)]


@pytest.mark.parametrize("order", [2, 3])
def test_signs_matching_shifted_and_gmres(order):
    from warp_preconditioners import linear
    from warp_preconditioners.block_ilu import BlockILU0
    from warp_preconditioners.mixed import MatchingSchur, MixedHarmonicSystem, ShiftedBlock

    n, alpha = 7, 0.04
    rng, k, mass = _graph(n)
    system = MixedHarmonicSystem(upload(k), vec(mass), order=order, data_weight=alpha)
    expected = _expected(k, mass, alpha, order)
    np.testing.assert_allclose(system.matrix.to_dense(), expected)
    H = k + alpha ** (1 / order) * np.diag(mass)
    np.testing.assert_allclose(system.H.to_dense(), H, rtol=1e-15, atol=1e-15)
    invH = np.linalg.inv(H)

    pre = MatchingSchur(system, max_row_size=n, kap_tolerance=0)
    r = rng.normal(size=(n, order))
    x = vec(r)
    y = torch.empty_like(x)
    pre.matvec(x, y, y, 1.0, 0.0)
    if order == 2:
        g = r[:, 0] - k @ (r[:, 1] / mass)
        u = invH @ np.diag(mass) @ invH @ g
        want = np.column_stack([u, -(k @ u + r[:, 1]) / mass])
    else:
        g = r[:, 0] - k @ (r[:, 1] / mass) + k @ ((k @ (r[:, 2] / mass)) / mass)
        u = -invH @ np.diag(mass) @ invH @ np.diag(mass) @ invH @ g
        a = -(k @ u + r[:, 2]) / mass
        want = np.column_stack([u, a, -(k @ a + r[:, 1]) / mass])
    np.testing.assert_allclose(host(y), want, rtol=2e-11, atol=2e-11)

    shifted = ShiftedBlock(system, max_row_size=n, kap_tolerance=0)
    shifted.matvec(x, y, y, 1.0, 0.0)
    if order == 2:
        s = alpha ** 0.25
        target = np.column_stack([invH @ ((r[:, 0] / s - s * r[:, 1]) * 0.5) / s,
                                  invH @ ((-r[:, 0] / s - s * r[:, 1]) * 0.5) * s])
    else:
        target = np.column_stack([-invH @ r[:, 2], -invH @ r[:, 1], -invH @ r[:, 0]])
    np.testing.assert_allclose(host(y), target, rtol=1e-11, atol=1e-11)

    # captured apply, replayed with eager work in between
    y.zero_()
    torch.cuda.synchronize()
    g_ = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g_):
        shifted.matvec(x, y, y, 1.0, 0.0)
    for _ in range(3):
        y.fill_(float("nan"))
        g_.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(y), target, rtol=1e-11, atol=1e-11)

    truth = rng.normal(size=(n, order))
    b = vec((expected @ truth.ravel()).reshape(n, order))
    solution = torch.zeros_like(b)
    linear.gmres(system.matrix, b, solution,
                 M=BlockILU0(system.matrix, factor_sweeps=20, solve_sweeps=n),
                 tol=1e-11, maxiter=60, restart=15)
    np.testing.assert_allclose(host(solution), truth, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("order", [2, 3])
def test_equilibration_preserves_equations(order):
    from warp_preconditioners import linear
    from warp_preconditioners.mixed import MixedHarmonicSystem, ShiftedBlock

    n = 5
    k = np.diag(np.full(n, 2.0)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)
    mass = np.linspace(0.6, 1.3, n)
    K, mw = upload(k), vec(mass)
    original = MixedHarmonicSystem(K, mw, order=order, data_weight=1e-4)
    balanced = MixedHarmonicSystem(K, mw, order=order, data_weight=1e-4, equilibrate=True)
    a = original.matrix.to_dense()
    s = np.tile(np.asarray(balanced.scale), n)
    np.testing.assert_allclose(balanced.matrix.to_dense(), s[:, None] * a * s[None, :],
                               rtol=1e-12, atol=1e-12)
    b = vec(np.linspace(-1, 1, n))
    rhs = balanced.rhs(b)
    np.testing.assert_allclose(host(rhs).ravel(), host(original.rhs(b)).ravel() * s)
    x = torch.zeros_like(rhs)
    linear.gmres(balanced.matrix, rhs, x, M=ShiftedBlock(balanced, max_row_size=n, kap_tolerance=0),
                 tol=1e-12, restart=15, maxiter=60)
    result = host(balanced.transform(x)).ravel()
    np.testing.assert_allclose(a @ result, host(original.rhs(b)).ravel(), atol=1e-9)
    np.testing.assert_array_equal(host(balanced.solution(x)), result.reshape(n, order)[:, 0])


def _grid_system(oracle, side, order, dtype, seed):
    rng = np.random.default_rng(seed)
    k = laplacian_2d(side) * (side + 1) ** 2 / 8.0
    n = k.shape[0]
    mass = rng.uniform(0.5, 1.5, size=n) / (side + 1) ** 2
    npd = np.float32 if dtype == torch.float32 else np.float64
    return rng, k, mass.astype(npd), npd


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_preconditioners_match_oracle(oracle, order, dtype):
    from warp_preconditioners.mixed import MatchingSchur, MixedHarmonicSystem, ShiftedBlock

    rng, k, mass, npd = _grid_system(oracle, 40, order, dtype, 5 + order)
    n = k.shape[0]
    alpha = 0.3
    system = MixedHarmonicSystem(upload(k, dtype), vec(mass, dtype), order=order,
                                 data_weight=alpha)
    ref = oracle.mixed_system(oracle.CSR.from_scipy(ss.csr_matrix(k), npd), mass, order=order,
                              data_weight=alpha)
    # identical assembled structure and values
    np.testing.assert_array_equal(host(system.matrix.offsets), ref["matrix"].off)
    np.testing.assert_array_equal(host(system.matrix.columns), ref["matrix"].col)
    np.testing.assert_array_equal(host(system.matrix.values), ref["matrix"].val)
    np.testing.assert_array_equal(host(system.H.values), ref["H"].val)
    P = oracle.fsai(ref["H"], max_row_size=8)
    inverse = lambda v: oracle.fsai_apply(P, np.ascontiguousarray(v, npd))  # noqa: E731
    x0 = rng.normal(size=(n, order)).astype(npd)
    y0 = rng.normal(size=(n, order)).astype(npd)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    for cls, apply in ((MatchingSchur, oracle.matching_schur_apply),
                       (ShiftedBlock, oracle.shifted_block_apply)):
        pre = cls(system, max_row_size=8)
        want = apply(ref, inverse, x0, y0, 0.8, -0.3)
        scale = np.abs(want).max()
        x, y = vec(x0, dtype), vec(y0, dtype)
        z = torch.empty_like(x)
        pre.matvec(x, y, z, 0.8, -0.3)
        np.testing.assert_allclose(host(z), want, rtol=0, atol=tol * scale)
        # z aliasing x, then y: the same values
        pre.matvec(x, y, x, 0.8, -0.3)
        np.testing.assert_array_equal(host(x), host(z))
        y2 = vec(y0, dtype)
        pre.matvec(vec(x0, dtype), y2, y2, 0.8, -0.3)
        np.testing.assert_array_equal(host(y2), host(z))
        # beta = 0 never reads y; alpha = 0 gives beta * y
        nan = torch.full_like(z, float("nan"))
        pre.matvec(vec(x0, dtype), nan, z, 0.8, 0.0)
        np.testing.assert_allclose(host(z), apply(ref, inverse, x0, None, 0.8, 0.0), rtol=0,
                                   atol=tol * scale)
        pre.matvec(torch.full_like(z, float("nan")), y, z, 0.0, 0.5)
        np.testing.assert_array_equal(host(z), (npd(0.5) * y0).astype(npd))


@pytest.mark.parametrize("cls", ["ShiftedBlock", "MatchingSchur"])
def test_gmres_with_block_preconditioners_converges(cls):
    # P1-scaled 2D mixed biharmonic (stiffness O(1), lumped mass h^2), 1,800
    # unknowns: right-preconditioned GMRES(30) with either block preconditioner
    # (FSAI W16 for H) reaches 1e-8, where unpreconditioned GMRES stalls near 1e-3
    # (scipy's GMRES on the oracle's operators: 123 / 207 iterations).
    from warp_preconditioners import linear, mixed

    side = 30
    k = laplacian_2d(side)
    n = k.shape[0]
    mass = np.full(n, 1.0 / (side + 1) ** 2)
    system = mixed.MixedHarmonicSystem(upload(k), vec(mass), order=2, data_weight=1.0)
    rng = np.random.default_rng(3)
    truth = rng.normal(size=(n, 2))
    A = system.matrix.to_scipy()
    b = vec((A @ truth.ravel()).reshape(n, 2))
    x = torch.zeros_like(b)
    pre = getattr(mixed, cls)(system, max_row_size=16)
    it, res, target = linear.gmres(system.matrix, b, x, M=pre, tol=1e-8, restart=30, maxiter=600)
    assert res <= target and it < 600
    np.testing.assert_allclose(host(x), truth, rtol=0, atol=1e-5)
    plain = torch.zeros_like(b)
    _, res_plain, _ = linear.gmres(system.matrix, b, plain, tol=1e-8, restart=30, maxiter=300)
    assert res_plain > 100 * target


def test_validation_messages():
    from warp_preconditioners.mixed import MatchingSchur, MixedHarmonicSystem

    _, k, mass = _graph(4)
    K = upload(k)
    with pytest.raises(ValueError, match="square scalar BSR"):
        MixedHarmonicSystem(upload(np.eye(4), block=2), vec(mass))
    with pytest.raises(ValueError, match="mass must match K"):
        MixedHarmonicSystem(K, vec(mass, torch.float32))
    with pytest.raises(ValueError, match="order must be 2 or 3"):
        MixedHarmonicSystem(K, vec(mass), order=4)
    with pytest.raises(ValueError, match="data_weight finite positive"):
        MixedHarmonicSystem(K, vec(mass), data_weight=-1.0)
    bad = mass.copy()
    bad[1] = 0.0
    with pytest.raises(ValueError, match="mass must be finite positive"):
        MixedHarmonicSystem(K, vec(bad))
    system = MixedHarmonicSystem(K, vec(mass))
    with pytest.raises(ValueError, match="b must match mass"):
        system.rhs(vec(np.ones(3)))
    with pytest.raises(ValueError, match="interleaved system vector"):
        system.solution(vec(np.ones((4, 3))))
    other = MixedHarmonicSystem(upload(k[:3, :3]), vec(mass[:3]))
    with pytest.raises(ValueError, match="H_inverse must match H"):
        MatchingSchur(system, H_inverse=MatchingSchur(other).inverse)
