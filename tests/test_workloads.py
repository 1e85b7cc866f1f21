"""The host workload generators pinned by closed forms and invariants: the jittered
Kuhn tet mesh, the P1 stiffness and lumped mass, the squared operator and the
K (x) C coupling (configurations 2-5 are built from them)."""

import numpy as np
import scipy.sparse as ss

from warp_preconditioners import workloads


def test_p1_element_matrix_of_the_reference_tet():
    # One tet (0,0,0), (1,0,0), (0,1,0), (0,0,1): volume 1/6, gradients
    # (-1,-1,-1), e1, e2, e3, so K_e = [[3,-1,-1,-1], [-1,1,0,0], ...] / 6 and the
    # lumped mass is 1/24 per vertex.
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], dtype=float)
    K, m = workloads.p1_laplacian(v, np.array([[0, 1, 2, 3]], dtype=np.int32))
    want = np.array([[3, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]]) / 6.0
    np.testing.assert_allclose(K.toarray(), want, rtol=0, atol=1e-15)
    np.testing.assert_allclose(m, np.full(4, 1 / 24), rtol=1e-15)


def test_p1_laplacian_invariants_on_jittered_mesh():
    # On the jittered 7^3 Kuhn mesh of the unit cube: the six tets per cube fill it
    # (total volume and lumped mass 1, all volumes positive); K is symmetric
    # positive semidefinite, annihilates constants, and annihilates every linear
    # field on interior rows (P1 reproduces linear functions); the Dirichlet energy
    # of a linear field c . x is |c|^2 times the volume, exactly up to rounding.
    v, t, boundary = workloads.kuhn_grid(7, 0.15, seed=3)
    assert len(t) == 6 * 6 ** 3
    vol, _ = workloads._gradients(v, t)
    e = v[t]
    signed = np.einsum("ti,ti->t", np.cross(e[:, 1] - e[:, 0], e[:, 2] - e[:, 0]), e[:, 3] - e[:, 0])
    assert (np.abs(signed) > 1e-6).all()
    np.testing.assert_allclose(vol.sum(), 1.0, rtol=1e-13)
    K, m = workloads.p1_laplacian(v, t)
    assert (m > 0).all()
    np.testing.assert_allclose(m.sum(), 1.0, rtol=1e-13)
    scale = abs(K).max()
    assert abs(K - K.T).max() <= 1e-15 * scale
    np.testing.assert_allclose(K @ np.ones(len(v)), 0.0, atol=1e-13 * scale)
    rng = np.random.default_rng(0)
    for _ in range(3):
        c, c0 = rng.normal(size=3), rng.normal()
        u = v @ c + c0
        np.testing.assert_allclose((K @ u)[~boundary], 0.0, atol=1e-12 * scale)
        np.testing.assert_allclose(u @ (K @ u), c @ c, rtol=1e-12)
    assert np.linalg.eigvalsh(K.toarray()).min() > -1e-12 * scale
    # a jittered interior vertex really moved (the mesh is not the regular grid)
    h = 1 / 6
    assert np.abs(v[~boundary] / h - np.round(v[~boundary] / h)).max() > 0.05


def test_squared_operator_invariants():
    # Q = alpha M + K M^-1 K: symmetric positive definite with Q 1 = alpha m (K 1 = 0)
    # and, for a linear field on a mesh, the energy u^T Q u = alpha u^T M u + |K u|^2_{M^-1}.
    v, t, _ = workloads.kuhn_grid(6, 0.15, seed=1)
    K, m = workloads.p1_laplacian(v, t)
    for alpha in (1.0, 0.05):
        Q = workloads.squared_operator(K, m, alpha)
        assert abs(Q - Q.T).max() <= 1e-14 * abs(Q).max()
        np.testing.assert_allclose(Q @ np.ones(len(v)), alpha * m, rtol=1e-9, atol=1e-12)
        u = np.random.default_rng(2).normal(size=len(v))
        Ku = K @ u
        np.testing.assert_allclose(u @ (Q @ u), alpha * (m * u) @ u + Ku @ (Ku / m), rtol=1e-12)
        assert np.linalg.eigvalsh(Q.toarray()).min() > 0
    # the pattern is K's two-ring: 65 entries per interior row of the Kuhn lattice
    v, t, boundary = workloads.kuhn_grid(8, 0.15, seed=5)
    K, m = workloads.p1_laplacian(v, t)
    Q = workloads.squared_operator(K, m)
    deep = np.all((np.indices((8, 8, 8)).reshape(3, -1).T > 1) &
                  (np.indices((8, 8, 8)).reshape(3, -1).T < 6), axis=1)
    assert set(np.diff(Q.indptr)[deep]) == {65}
    assert set(np.diff(K.indptr)[~boundary]) == {15}


def test_kron_coupling_spectrum():
    # K (x) C with C = R diag(1, 1, contrast) R^T, R orthogonal: the eigenvalues are
    # products of the 7-point Laplacian's (closed form) and C's.
    n = 5
    off, col, val, c, k = workloads.kron_coupling(n, 1e3, seed=4)
    r = workloads.haar_rotation(4)
    np.testing.assert_allclose(r @ r.T, np.eye(3), atol=1e-14)
    np.testing.assert_allclose(np.linalg.eigvalsh(c), [1, 1, 1e3], rtol=1e-12)
    a = ss.bsr_matrix((val, col, off), shape=(3 * n ** 3, 3 * n ** 3)).toarray()
    lam1 = 2 - 2 * np.cos(np.arange(1, n + 1) * np.pi / (n + 1))
    lam = (lam1[:, None, None] + lam1[None, :, None] + lam1[None, None, :]).ravel()
    want = np.sort(np.concatenate([lam, lam, 1e3 * lam]))
    np.testing.assert_allclose(np.linalg.eigvalsh(a), want, rtol=1e-10, atol=1e-9)
