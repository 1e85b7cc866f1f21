"""GPU parity for FSAI setup and apply (reference tests/test_fsai.py, test_fsai_float16.py),
plus bit-exact pattern/value parity of the HIP build against the CPU oracle."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dense, dev, host, laplacian_2d, oracle_csr, random_spd, upload, vec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wp():
    import warp_preconditioners as w

    return w


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("step", [1, 3])
def test_exact_full_pattern(wp, dtype, step):
    rng = np.random.default_rng(7)
    a = rng.normal(size=(8, 8))
    a = a @ a.T + np.eye(8)
    m = wp.FSAI(upload(a, dtype), max_row_size=8, kap_tolerance=0, max_step_size=step)
    g = dense(m.G)
    np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=1e-5, atol=1e-6)
    assert np.allclose(g, np.tril(g))
    x = vec(rng.normal(size=8), dtype)
    y = vec(rng.normal(size=8), dtype)
    expected = 1.7 * g.T @ g @ host(x) - 0.3 * host(y)
    m.matvec(x, y, y, 1.7, -0.3)
    np.testing.assert_allclose(host(y), expected, rtol=1e-5, atol=1e-6)


def test_jacobi_blocks_and_cg(wp):
    from warp_preconditioners import linear

    a = np.diag(np.full(12, 3.0)) + np.diag(np.full(11, -1.0), 1) + np.diag(np.full(11, -1.0), -1)
    A = upload(a, block=3)
    m = wp.FSAI(A, max_row_size=1)
    np.testing.assert_allclose(dense(m.G), np.eye(12) / np.sqrt(3))
    m = wp.FSAI(A, max_row_size=8)
    b = torch.ones((4, 3), dtype=torch.float64, device=dev())
    x = torch.zeros_like(b)
    linear.cg(A, b, x, M=m, tol=1e-12, maxiter=100)
    np.testing.assert_allclose(a @ host(x).ravel(), 1.0, atol=1e-10)


def test_invalid_diagonal(wp):
    with pytest.raises(ValueError, match="positive diagonals"):
        wp.FSAI(upload(np.diag([1.0, 0.0])))


def test_sparse_reference_and_aliases(wp):
    rng = np.random.default_rng(18)
    a = np.diag(np.full(20, 4.0)) + np.diag(np.full(19, -1.0), 1) + np.diag(np.full(19, -1.0), -1)
    scale = np.geomspace(1e-4, 1e4, 20)
    a = scale[:, None] * a * scale[None, :]
    m = wp.FSAI(upload(a), max_row_size=5, kap_tolerance=0)
    g = dense(m.G)
    assert np.linalg.eigvalsh(g.T @ g).min() > 0
    assert np.count_nonzero(g, axis=1).max() <= 5
    for i in range(20):
        support = np.flatnonzero(g[i])
        e = np.zeros(len(support))
        e[-1] = 1
        ref = np.linalg.solve(a[np.ix_(support, support)], e)
        ref /= np.sqrt(ref[-1])
        np.testing.assert_allclose(g[i, support], ref, rtol=1e-10, atol=1e-12)
    for alias in ["none", "xz", "yz", "xy", "xyz"]:
        x = vec(rng.normal(size=20))
        y = x if alias in ["xy", "xyz"] else vec(rng.normal(size=20))
        z = x if alias in ["xz", "xyz"] else y if alias == "yz" else torch.empty_like(x)
        expected = 0.7 * g.T @ g @ host(x) - 1.2 * host(y)
        m.matvec(x, y, z, 0.7, -1.2)
        np.testing.assert_allclose(host(z), expected, rtol=1e-12, atol=1e-10)
    x = vec(rng.normal(size=20))
    y = torch.full_like(x, float("nan"))
    m.matvec(x, y, y, 1.0, 0.0)
    assert np.isfinite(host(y)).all()
    # Captured apply replayed several times, eager work in between.
    z = torch.empty_like(x)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        m.matvec(x, y, z, 1.0, 0.0)
    for _ in range(3):
        z.fill_(0.0)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_allclose(host(z), g.T @ g @ host(x), rtol=1e-12, atol=1e-10)


def test_small_pivot_retains_spd_factor(wp):
    a = np.array([[1.0, 1.0 - 1e-14], [1.0 - 1e-14, 1.0]])
    m = wp.FSAI(upload(a), pivot_floor=1e-12)
    assert m.truncated_rows == 1
    g = dense(m.G)
    assert np.isfinite(g).all() and np.linalg.eigvalsh(g.T @ g).min() > 0


def test_batch_width_cap_and_pivot_guard(wp):
    a = np.array([[1.0, 1.0 - 1e-14, 0.0], [1.0 - 1e-14, 1.0, 0.0], [0.0, 0.0, 2.0]])
    pre = wp.FSAI(upload(a), max_row_size=2, max_step_size=6)
    assert pre.truncated_rows == 1
    np.testing.assert_allclose(dense(pre.G), np.diag(1 / np.sqrt(np.diag(a))))
    pre = wp.FSAI(upload(a), max_row_size=1, max_step_size=6)
    np.testing.assert_allclose(dense(pre.G), np.diag(1 / np.sqrt(np.diag(a))))


def test_invalid_step_size(wp):
    a = upload(np.eye(3))
    for value in [0, -1, 65, 1.5]:
        with pytest.raises(ValueError, match="max_step_size"):
            wp.FSAI(a, max_step_size=value)


@pytest.mark.parametrize("step", [2, 3])
def test_batch_grows_through_narrow_frontier(wp, step):
    a = np.diag(np.full(12, 3.0)) + np.diag(-np.ones(11), 1) + np.diag(-np.ones(11), -1)
    pre = wp.FSAI(upload(a), max_row_size=8, max_step_size=step, kap_tolerance=0.0)
    g = dense(pre.G)
    np.testing.assert_array_equal(np.count_nonzero(g, axis=1), np.minimum(np.arange(1, 13), 8))
    for i in range(12):
        support = np.arange(max(0, i - 7), i + 1)
        rhs = np.zeros(len(support))
        rhs[-1] = 1
        z = np.linalg.solve(a[np.ix_(support, support)], rhs)
        np.testing.assert_allclose(g[i, support], z / np.sqrt(z[-1]), rtol=1e-12)


def test_zero_tolerance_disables_roundoff_stopping(wp):
    n = 32
    a = np.diag(np.full(n, 3.0)) + np.diag(-np.ones(n - 1), 1) + np.diag(-np.ones(n - 1), -1)
    pre = wp.FSAI(upload(a), max_row_size=n, kap_tolerance=0, max_step_size=3)
    np.testing.assert_array_equal(np.diff(host(pre.G.offsets)), np.arange(1, n + 1))
    g = dense(pre.G)
    np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=1e-12, atol=1e-15)


def _assert_same_factor(gpu, orc, exact=True):
    np.testing.assert_array_equal(host(gpu.G.offsets), orc["G"].off)
    np.testing.assert_array_equal(host(gpu.G.columns), orc["G"].col)
    np.testing.assert_array_equal(host(gpu.GT.offsets), orc["GT"].off)
    np.testing.assert_array_equal(host(gpu.GT.columns), orc["GT"].col)
    gv = host(gpu.G.values).astype(np.float64)
    ov = orc["G"].val.astype(np.float64)
    if exact:
        np.testing.assert_array_equal(gv, ov)
    else:
        np.testing.assert_allclose(gv, ov, rtol=1e-5, atol=0)
    assert gpu.truncated_rows == orc["truncated_rows"]


@pytest.mark.parametrize("dtype", [torch.float16, torch.float32, torch.float64])
@pytest.mark.parametrize("width,step,tol", [(4, 1, 1e-3), (8, 1, 0.0), (8, 3, 1e-3),
                                            (16, 2, 0.0), (33, 5, 0.0), (64, 1, 0.0)])
def test_build_matches_oracle_bitwise(wp, oracle, dtype, width, step, tol):
    rng = np.random.default_rng(width * 7 + step)
    n = 160
    a = random_spd(rng, n, 0.06, (0.3, 3.0) if dtype == torch.float16 else (1e-2, 1e2))
    npd = {torch.float16: np.float16, torch.float32: np.float32, torch.float64: np.float64}[dtype]
    a = a.astype(npd).astype(np.float64)
    A = upload(a, dtype)
    pre = wp.FSAI(A, max_row_size=width, max_step_size=step, kap_tolerance=tol)
    orc = oracle.fsai(oracle_csr(oracle, a, npd), max_row_size=width, max_step_size=step,
                      kap_tolerance=tol)
    _assert_same_factor(pre, orc)


@pytest.mark.parametrize("block", [1, 3])
def test_build_matches_oracle_on_blocks_and_grid(wp, oracle, block):
    # 2D five-point Laplacian (24x24 interior grid) and its
    # 3x3-blocked copy: the factor pattern and values must match bit for bit.
    a = laplacian_2d(24).toarray() if block == 1 else np.kron(laplacian_2d(8).toarray(),
                                                               np.array([[4, 1, 0], [1, 4, 1], [0, 1, 4.0]]))
    A = upload(a, torch.float32, block)
    pre = wp.FSAI(A, max_row_size=8)
    orc = oracle.fsai(oracle_csr(oracle, a, np.float32, block), max_row_size=8)
    _assert_same_factor(pre, orc)


def test_factor_conversion_and_mixed_cg(wp, oracle):
    from warp_preconditioners import linear

    rng = np.random.default_rng(72)
    n = 25
    a = np.diag(np.full(n, 3.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    A = upload(a)
    pre = wp.FSAI(A, max_row_size=8, factor_dtype=torch.float32, apply_lanes=4)
    orc = oracle.fsai(oracle_csr(oracle, a), max_row_size=8, factor_dtype=np.float32)
    _assert_same_factor(pre, orc)
    g = dense(pre.G)
    np.testing.assert_array_equal(dense(pre.GT), g.T)
    x = vec(rng.normal(size=n))
    y = vec(rng.normal(size=n))
    expected = 0.7 * g.T @ g @ host(x) - 0.3 * host(y)
    pre.matvec(x, y, x, 0.7, -0.3)
    np.testing.assert_allclose(host(x), expected, rtol=1e-12, atol=1e-12)
    b = torch.ones(n, dtype=torch.float64, device=dev())
    x.zero_()
    linear.cg(A, b, x, M=pre, tol=1e-12, maxiter=200)
    np.testing.assert_allclose(a @ host(x), 1.0, atol=1e-11)
    with pytest.raises(ValueError, match="positive diagonal"):
        wp.FSAI(upload(np.eye(3) * 1e100), factor_dtype=torch.float32)


def test_half_factor_application(wp):
    rng = np.random.default_rng(19)
    r = rng.normal(size=(9, 9))
    a = (r @ r.T + 4 * np.eye(9)).astype(np.float16).astype(np.float64)
    raw = upload(a, torch.float16, 3)
    for lanes in [1, 4]:
        pre = wp.FSAI(raw, max_row_size=9, kap_tolerance=0, apply_lanes=lanes)
        assert pre.G.scalar_type == torch.float16 and pre.GT.scalar_type == torch.float16
        g = dense(pre.G)
        np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=1e-2, atol=1e-3)
        x = torch.tensor(rng.normal(size=(3, 3)), dtype=torch.float16, device=dev())
        xv = host(x).astype(np.float64).ravel()
        expected = 0.7 * g.T @ g @ xv - 0.2 * xv
        pre.matvec(x, x, x, 0.7, -0.2)
        np.testing.assert_allclose(host(x).astype(np.float64).ravel(), expected, rtol=5e-3, atol=1e-3)
    raw = upload(np.diag([1, 0.01]), torch.float16, 2)
    pre = wp.FSAI(raw)
    x = torch.ones(2, dtype=torch.float16, device=dev())
    y = torch.zeros_like(x)
    pre.matvec(x, y, y, 1, 0)
    np.testing.assert_allclose(host(y).astype(np.float64), [1, 100], rtol=0, atol=0.01)
    for a in [np.array([[1, 2], [2, 4.0]]), np.array([[9, 21], [21, 49.0]])]:
        pre = wp.FSAI(upload(a, torch.float16), max_row_size=2)
        assert pre.truncated_rows == 1
        np.testing.assert_allclose(dense(pre.G), np.diag(1 / np.sqrt(np.diag(a))), rtol=2e-3)


@pytest.mark.parametrize("factor", [torch.float16, torch.bfloat16])
@pytest.mark.parametrize("vdt", [torch.float32, torch.float64])
def test_row_scaled_narrow_factor(wp, factor, vdt):
    # Row-scaled fp16/bf16 storage: G^T G x within the storage precision of the
    # full-precision factor.
    rng = np.random.default_rng(3)
    a = random_spd(rng, 300, 0.03, (1e-3, 1e3))
    A = upload(a)
    ref = wp.FSAI(A, max_row_size=8)
    pre = wp.FSAI(A, max_row_size=8, factor_dtype=factor)
    g = dense(ref.G)
    gq = pre.dense_factor()
    np.testing.assert_array_equal(host(pre.G.columns), host(ref.G.columns))
    rel = 4e-3 if factor == torch.float16 else 2e-2
    row_max = np.abs(g).max(axis=1, keepdims=True)
    assert np.all(np.abs(gq - g) <= rel * row_max)
    x = torch.tensor(rng.normal(size=300), dtype=vdt, device=dev())
    z = torch.empty_like(x)
    pre.matvec(x, x, z, 1.0, 0.0)
    expected = gq.T @ gq @ host(x).astype(np.float64)
    tol = 1e-5 if vdt == torch.float32 else 1e-12
    np.testing.assert_allclose(host(z), expected, rtol=tol, atol=tol * np.abs(expected).max())


@pytest.mark.parametrize("nvec", [2, 3, 4, 8])
def test_batched_apply_equals_single(wp, nvec):
    rng = np.random.default_rng(11)
    a = random_spd(rng, 500, 0.02)
    pre = wp.FSAI(upload(a), max_row_size=8, apply_lanes=8)
    X = torch.tensor(rng.normal(size=(500, nvec)), dtype=torch.float64, device=dev())
    Z = torch.empty_like(X)
    pre.apply_batch(X, X, Z, 1.3, 0.2, nvec)
    for v in range(nvec):
        x = X[:, v].contiguous()
        z = torch.empty_like(x)
        pre.matvec(x, x, z, 1.3, 0.2)
        np.testing.assert_allclose(host(Z[:, v]), host(z), rtol=1e-13, atol=1e-13)


def test_empty_matrix(wp):
    from warp_preconditioners.sparse import bsr_zeros

    for dtype in [torch.float32, torch.float64]:
        pre = wp.FSAI(bsr_zeros(0, 0, dtype, dev()), reuse_pattern=True)
        assert pre.update() is pre
        assert pre.G.nnz_sync() == 0 and pre.GT.nnz_sync() == 0
        v = torch.empty(0, dtype=dtype, device=dev())
        pre.matvec(v, v, v, 1.0, 0.0)
        torch.cuda.synchronize()


@pytest.mark.parametrize("count", [1, 1023, 1024, 8191, 8192, 8193, 100003, 262145])
def test_scan_inclusive_exact(count):
    # The setup scan (tiles of 8192: thread-serial runs, wavefront and workgroup
    # carries) equals numpy's cumulative sum on counts across tile boundaries.
    from warp_preconditioners import _native as nat

    rng = np.random.default_rng(count)
    a = rng.integers(0, 70, size=count).astype(np.int32)
    src = torch.tensor(a, device=dev())
    out = torch.full((count,), -1, dtype=torch.int32, device=dev())
    nat.call("wpc_scan_inclusive", count, nat.ptr(src), nat.ptr(out), nat.stream())
    np.testing.assert_array_equal(host(out), np.cumsum(a, dtype=np.int64).astype(np.int32))
