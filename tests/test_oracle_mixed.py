"""The oracle's mixed harmonic systems (mixed.py) pinned by closed forms.

The checks follow tests/test_indefinite.py:58-139 and :195-229: the interleaved
matrix equals the permuted dense block matrix with the benchmark's signs, its
Schur complement is alpha M + K M^-1 K (order 2) or alpha M - K M^-1 K M^-1 K
(order 3), the matching-Schur and shifted-block applies with an exact H inverse
(full-width FSAI, kap_tolerance 0) equal their dense formulas, and equilibration
is the congruence D A D.
"""

import numpy as np
import pytest
import scipy.sparse as ss


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


def _exact_inverse(oracle, system, n):
    P = oracle.fsai(system["H"], max_row_size=n, kap_tolerance=0)
    return lambda v: oracle.fsai_apply(P, np.ascontiguousarray(v))


@pytest.mark.parametrize("order", [2, 3])
def test_signs_schur_and_preconditioners(oracle, order):
    n, alpha = 7, 0.04
    rng, k, mass = _graph(n)
    K = oracle.CSR.from_scipy(ss.csr_matrix(k))
    system = oracle.mixed_system(K, mass, order=order, data_weight=alpha)
    expected = _expected(k, mass, alpha, order)
    np.testing.assert_allclose(system["matrix"].to_dense(), expected)
    primary = np.arange(0, n * order, order)
    aux = np.setdiff1d(np.arange(n * order), primary)
    schur = expected[np.ix_(primary, primary)] - expected[np.ix_(primary, aux)] @ np.linalg.solve(
        expected[np.ix_(aux, aux)], expected[np.ix_(aux, primary)])
    im = np.diag(1 / mass)
    target = alpha * np.diag(mass) + (k @ im @ k if order == 2 else -k @ im @ k @ im @ k)
    np.testing.assert_allclose(schur, target, rtol=1e-12, atol=1e-12)
    H = k + alpha ** (1 / order) * np.diag(mass)
    np.testing.assert_allclose(system["H"].to_dense(), H, rtol=1e-15, atol=1e-15)

    inverse = _exact_inverse(oracle, system, n)
    invH = np.linalg.inv(H)
    r = rng.normal(size=(n, order))
    got = oracle.matching_schur_apply(system, inverse, r)
    if order == 2:
        g = r[:, 0] - k @ (r[:, 1] / mass)
        u = invH @ np.diag(mass) @ invH @ g
        want = np.column_stack([u, -(k @ u + r[:, 1]) / mass])
    else:
        g = r[:, 0] - k @ (r[:, 1] / mass) + k @ ((k @ (r[:, 2] / mass)) / mass)
        u = -invH @ np.diag(mass) @ invH @ np.diag(mass) @ invH @ g
        a = -(k @ u + r[:, 2]) / mass
        want = np.column_stack([u, a, -(k @ a + r[:, 1]) / mass])
    np.testing.assert_allclose(got, want, rtol=2e-11, atol=2e-11)

    got = oracle.shifted_block_apply(system, inverse, r)
    if order == 2:
        s = alpha ** 0.25
        want = np.column_stack([invH @ ((r[:, 0] / s - s * r[:, 1]) * 0.5) / s,
                                invH @ ((-r[:, 0] / s - s * r[:, 1]) * 0.5) * s])
    else:
        want = np.column_stack([-invH @ r[:, 2], -invH @ r[:, 1], -invH @ r[:, 0]])
    np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("order", [2, 3])
def test_equilibration_is_a_congruence(oracle, order):
    n = 5
    k = np.diag(np.full(n, 2.0)) - np.diag(np.ones(n - 1), 1) - np.diag(np.ones(n - 1), -1)
    mass = np.linspace(0.6, 1.3, n)
    K = oracle.CSR.from_scipy(ss.csr_matrix(k))
    original = oracle.mixed_system(K, mass, order=order, data_weight=1e-4)
    balanced = oracle.mixed_system(K, mass, order=order, data_weight=1e-4, equilibrate=True)
    a = original["matrix"].to_dense()
    s = np.tile(balanced["scale"], n)
    np.testing.assert_allclose(balanced["matrix"].to_dense(), s[:, None] * a * s[None, :],
                               rtol=1e-12, atol=1e-12)
    assert balanced["data_weight"] == 1.0


def test_alpha_beta_semantics(oracle):
    n = 6
    rng, k, mass = _graph(n, seed=11)
    K = oracle.CSR.from_scipy(ss.csr_matrix(k))
    system = oracle.mixed_system(K, mass, order=3, data_weight=0.5)
    inverse = _exact_inverse(oracle, system, n)
    x, y = rng.normal(size=(2, n, 3))
    for apply in (oracle.matching_schur_apply, oracle.shifted_block_apply):
        base = apply(system, inverse, x)
        np.testing.assert_allclose(apply(system, inverse, x, y, 0.7, -0.2), 0.7 * base - 0.2 * y,
                                   rtol=1e-14, atol=1e-14)
        np.testing.assert_array_equal(apply(system, inverse, x, y, 0.0, 0.5), 0.5 * y)


def test_nonpositive_mass_rejected(oracle):
    _, k, mass = _graph(4)
    mass[2] = 0.0
    with pytest.raises(ValueError, match="finite positive"):
        oracle.mixed_system(oracle.CSR.from_scipy(ss.csr_matrix(k)), mass)
