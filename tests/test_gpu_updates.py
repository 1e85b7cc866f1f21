"""GPU parity for fixed-pattern refits and block updates (reference tests/test_updates.py)."""

import numpy as np
import pytest
import torch

from gpu_helpers import dense, dev, host, oracle_csr, random_spd, upload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wp():
    import warp_preconditioners as w

    return w


@pytest.mark.parametrize("block", [1, 3])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("step", [1, 3])
def test_refit_local_solves_and_rollback(wp, block, dtype, step):
    rng = np.random.default_rng(48)
    r = rng.normal(size=(12, 12))
    a = r @ r.T + np.eye(12) * 3
    A = upload(a, dtype, block)
    m = wp.FSAI(A, max_row_size=5, reuse_pattern=True, factor_dtype=torch.float32,
                max_step_size=step)
    gptr, tptr = m.G.values.data_ptr(), m.GT.values.data_ptr()
    b = a + np.diag(np.arange(12))
    B = upload(b, dtype, block)
    m.update(B)
    assert (m.G.values.data_ptr(), m.GT.values.data_ptr()) == (gptr, tptr)
    g = dense(m.G)
    np.testing.assert_array_equal(dense(m.GT), g.T)
    off, col = host(m.G.offsets), host(m.G.columns)
    for i in range(12):
        support = col[off[i]:off[i + 1]]
        e = np.zeros(len(support))
        e[-1] = 1
        z = np.linalg.solve(b[np.ix_(support, support)], e)
        np.testing.assert_allclose(g[i, support], z / np.sqrt(z[-1]), rtol=2e-4, atol=2e-7)
    from warp_preconditioners.sparse import bsr_copy

    invalid = bsr_copy(B)
    invalid.values.zero_()
    with pytest.raises(ValueError, match="positive diagonals"):
        m.update(invalid)
    np.testing.assert_array_equal(dense(m.G), g)
    changed = bsr_copy(B)
    changed.columns[0] += 1
    with pytest.raises(ValueError, match="sparsity"):
        m.update(changed)
    np.testing.assert_array_equal(dense(m.G), g)
    B.values.copy_(A.values)
    m.update()
    fresh = wp.FSAI(A, max_row_size=12, kap_tolerance=0, reuse_pattern=True)
    fresh.update(upload(b, dtype, block))
    np.testing.assert_allclose(dense(fresh.G).T @ dense(fresh.G), np.linalg.inv(b), rtol=2e-4,
                               atol=1e-7)
    assert fresh.quality(probes=2) < (1e-5 if dtype == torch.float32 else 1e-12)


def test_refit_matches_oracle_bitwise(wp, oracle):
    rng = np.random.default_rng(9)
    a = random_spd(rng, 200, 0.05, (1e-2, 1e2))
    b = a + np.diag(rng.uniform(0, 1, 200)) * np.diag(a)
    for dtype, npd in [(torch.float64, np.float64), (torch.float32, np.float32)]:
        m = wp.FSAI(upload(a, dtype), max_row_size=10, reuse_pattern=True)
        m.update(upload(b, dtype))
        P = oracle.fsai(oracle_csr(oracle, a, npd), max_row_size=10)
        Q = oracle.refit(P, oracle_csr(oracle, b, npd))
        np.testing.assert_array_equal(host(m.G.values), Q["G"].val)
        np.testing.assert_array_equal(host(m.GT.values), Q["GT"].val)


def test_block_update_and_capture(wp):
    a = np.array([[3.0, 0.4, 0], [0.4, 2, 0], [0, 0, 4.0]])
    A = upload(a, block=3)
    B = upload(2 * a, block=3)
    for m in [wp.BlockJacobi(A), wp.FSAI(A, max_row_size=3, kap_tolerance=0, reuse_pattern=True)]:
        x = torch.tensor([[1.0, 2.0, 3.0]], dtype=torch.float64, device=dev())
        y = torch.zeros_like(x)
        m.matvec(x, y, y, 1.0, 0.0)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            m.matvec(x, y, y, 1.0, 0.0)
        m.update(B)
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(y).ravel(), np.linalg.solve(2 * a, host(x).ravel()),
                                   rtol=1e-12)
        B.values.zero_()
        with pytest.raises(ValueError):
            m.update(B)
        m.matvec(x, y, y, 1.0, 0.0)
        np.testing.assert_allclose(host(y).ravel(), np.linalg.solve(2 * a, host(x).ravel()),
                                   rtol=1e-12)
        B.values.copy_(2 * A.values)


def test_numeric_zero_becomes_nonzero_inside_block(wp):
    a = np.array([[4.0, 1.0, 0.0], [1.0, 4.0, 1.0], [0.0, 1.0, 4.0]])
    b = a + 0.2 * np.ones((3, 3))
    A = upload(a, block=3)
    m = wp.FSAI(A, max_row_size=3, kap_tolerance=0, reuse_pattern=True)
    m.update(A)
    np.testing.assert_allclose(dense(m.G).T @ dense(m.G), np.linalg.inv(a), atol=1e-13)
    m.update(upload(b, block=3))
    np.testing.assert_allclose(dense(m.G).T @ dense(m.G), np.linalg.inv(b), atol=1e-13)


def test_refit_requires_explicit_plan(wp):
    a = upload(np.eye(3))
    with pytest.raises(ValueError, match="reuse_pattern=True"):
        wp.FSAI(a).update(a)


def test_half_refit_and_rollback(wp):
    rng = np.random.default_rng(19)
    r = rng.normal(size=(9, 9))
    a = (r @ r.T + 4 * np.eye(9)).astype(np.float16).astype(np.float64)
    pre = wp.FSAI(upload(a, torch.float16, 3), max_row_size=9, kap_tolerance=0, reuse_pattern=True)
    pre.update(upload(2 * a, torch.float16, 3))
    g = dense(pre.G)
    np.testing.assert_allclose(g.T @ g, np.linalg.inv(2 * a), rtol=1e-2, atol=1e-3)
    from warp_preconditioners.sparse import bsr_copy

    invalid = bsr_copy(upload(a, torch.float16, 3))
    invalid.values.zero_()
    with pytest.raises(ValueError):
        pre.update(invalid)
    np.testing.assert_array_equal(dense(pre.G), g)
    np.testing.assert_array_equal(dense(pre.GT), g.T)


@pytest.mark.parametrize("factor", [torch.bfloat16, torch.float16])
def test_scaled_factor_refit(wp, factor):
    rng = np.random.default_rng(5)
    a = random_spd(rng, 120, 0.05)
    b = a * 1.7 + np.diag(np.diag(a)) * 0.3
    pre = wp.FSAI(upload(a), max_row_size=8, reuse_pattern=True, factor_dtype=factor)
    pre.update(upload(b))
    full = wp.FSAI(upload(a), max_row_size=8, reuse_pattern=True)
    full.update(upload(b))
    g = dense(full.G)
    rel = 4e-3 if factor == torch.float16 else 2e-2
    assert np.all(np.abs(pre.dense_factor() - g) <= rel * np.abs(g).max(axis=1, keepdims=True))


def _probe(n, seed, dtype):
    from warp_preconditioners import _native as nat

    x = torch.empty(n, dtype=dtype, device=dev())
    nat.call("wpc_rademacher", n, seed, nat.ptr(x), nat.code(dtype), nat.stream())
    return host(x).astype(np.float64)


@pytest.mark.parametrize("factor_dtype", [None, torch.bfloat16])
def test_quality_of_inexact_factor_against_dense_estimator(wp, factor_dtype):
    # quality() on an inexact factor (width 4 on a 300-row SPD matrix): equal to the
    # dense evaluation of the same estimator, sqrt(sum_p |(G A G^T - I) x_p|^2 /
    # (probes n)) with the device's own probe vectors x_p (seed, seed + 1, ...), so
    # a wrong normalisation, sign or transpose fails; the probes are +-1 and differ
    # per seed; and over many probes the estimate approaches the exact
    # ||G A G^T - I||_F / sqrt(n) (its expectation).
    rng = np.random.default_rng(77)
    n = 300
    a = random_spd(rng, n, 0.03, (0.1, 10.0))
    A = upload(a)
    P = wp.FSAI(A, max_row_size=4, factor_dtype=factor_dtype)
    g = P.dense_factor()
    m = g @ a @ g.T - np.eye(n)
    exact = np.linalg.norm(m) / np.sqrt(n)
    assert exact > 1e-2  # an inexact factor
    probes, seed = 5, 23
    xs = [_probe(n, seed + p, torch.float64) for p in range(probes)]
    for x in xs:
        assert set(np.unique(x)) <= {-1.0, 1.0} and abs(x.sum()) < n
    assert not np.array_equal(xs[0], xs[1])
    dense_est = np.sqrt(sum(float(np.sum((m @ x) ** 2)) for x in xs) / (probes * n))
    got = P.quality(probes=probes, seed=seed)
    np.testing.assert_allclose(got, dense_est, rtol=1e-6 if factor_dtype else 1e-10)
    # a transposed product (G^T A G) or a missing -I gives a different value
    wrong_t = g.T @ a @ g - np.eye(n)
    assert abs(np.sqrt(sum(float(np.sum((wrong_t @ x) ** 2)) for x in xs) / (probes * n))
               - got) > 1e-3 * got
    many = P.quality(probes=400, seed=1)
    assert abs(many - exact) < 0.1 * exact
