"""Pins the CPU oracle's FSAI pipeline (fsai.py:12-344) with closed forms and an
independent numpy restatement of the greedy search.  The cases follow the
reference's own tests (tests/test_fsai.py, tests/test_fsai_float16.py)."""

import numpy as np
import pytest
import scipy.sparse as ss


def csr(cpu, a, dtype=np.float64, block=1):
    return cpu.CSR.from_scipy(ss.csr_matrix(a), dtype, block)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("step", [1, 3])
def test_exact_full_pattern(oracle, dtype, step):
    # tests/test_fsai.py:36-51: a full support reproduces the inverse.
    rng = np.random.default_rng(7)
    a = rng.normal(size=(8, 8))
    a = a @ a.T + np.eye(8)
    P = oracle.fsai(csr(oracle, a, dtype), max_row_size=8, kap_tolerance=0, max_step_size=step)
    g = P["G"].to_dense()
    np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=1e-5, atol=1e-6)
    assert np.allclose(g, np.tril(g))
    x, y = rng.normal(size=8).astype(dtype), rng.normal(size=8).astype(dtype)
    z = oracle.fsai_apply(P, x, y, 1.7, -0.3)
    np.testing.assert_allclose(z, 1.7 * g.T @ g @ x - 0.3 * y, rtol=1e-5, atol=1e-6)


def test_jacobi_width_one_on_blocks(oracle):
    # tests/test_fsai.py:54-59: width 1 is Jacobi exactly.
    a = np.diag(np.full(12, 3.0)) + np.diag(np.full(11, -1.0), 1) + np.diag(np.full(11, -1.0), -1)
    P = oracle.fsai(csr(oracle, a, block=3), max_row_size=1)
    np.testing.assert_allclose(P["G"].to_dense(), np.eye(12) / np.sqrt(3))


def test_invalid_diagonal(oracle):
    with pytest.raises(ValueError, match="positive diagonals"):
        oracle.fsai(csr(oracle, np.diag([1.0, 0.0])))


def test_local_solves_under_severe_scaling(oracle):
    # tests/test_fsai.py:73-90: every row solves its principal submatrix.
    a = np.diag(np.full(20, 4.0)) + np.diag(np.full(19, -1.0), 1) + np.diag(np.full(19, -1.0), -1)
    scale = np.geomspace(1e-4, 1e4, 20)
    a = scale[:, None] * a * scale[None, :]
    P = oracle.fsai(csr(oracle, a), max_row_size=5, kap_tolerance=0)
    g = P["G"].to_dense()
    assert np.linalg.eigvalsh(g.T @ g).min() > 0
    assert np.count_nonzero(g, axis=1).max() <= 5
    for i in range(20):
        support = np.flatnonzero(g[i])
        e = np.zeros(len(support))
        e[-1] = 1
        ref = np.linalg.solve(a[np.ix_(support, support)], e)
        ref /= np.sqrt(ref[-1])
        np.testing.assert_allclose(g[i, support], ref, rtol=1e-10, atol=1e-12)


def test_small_pivot_retains_spd_factor(oracle):
    a = np.array([[1.0, 1.0 - 1e-14], [1.0 - 1e-14, 1.0]])
    P = oracle.fsai(csr(oracle, a), pivot_floor=1e-12)
    assert P["truncated_rows"] == 1
    g = P["G"].to_dense()
    assert np.isfinite(g).all() and np.linalg.eigvalsh(g.T @ g).min() > 0


def greedy_reference(a, width, step, tolerance):
    """Independent restatement of the adaptive search (tests/test_fsai.py:143-169)."""
    equilibrated = a / np.sqrt(np.diag(a)[:, None] * np.diag(a)[None, :])
    rows = []
    for i in range(len(a)):
        support = [i]
        z = np.ones(1)
        while len(support) < width:
            frontier = sorted(set(np.flatnonzero(np.any(a[support] != 0, axis=0))) - set(support))
            frontier = [c for c in frontier if c < i]
            if not frontier:
                break
            scores = [abs(equilibrated[c, support] @ z) for c in frontier]
            ordered = sorted(zip(scores, frontier), key=lambda item: (-item[0], item[1]))
            selected = [c for s, c in ordered if s > 0][: min(step, width - len(support))]
            if not selected:
                break
            support.extend(selected)
            old = z[0]
            rhs = np.zeros(len(support))
            rhs[0] = 1
            z = np.linalg.solve(equilibrated[np.ix_(support, support)], rhs)
            if 1 - old / z[0] <= tolerance:
                break
        row = np.zeros(len(a))
        row[support] = z / np.sqrt(z[0] * np.diag(a)[support])
        rows.append(row)
    return np.array(rows)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("step", [1, 3])
def test_adaptive_support_matches_independent_greedy_search(oracle, dtype, step):
    # tests/test_fsai.py:125-169, stored zeros inside 3x3 blocks included.
    rng = np.random.default_rng(193)
    a = rng.normal(size=(18, 18))
    a[rng.random(a.shape) < 0.65] = 0
    a = a + a.T
    np.fill_diagonal(a, np.sum(np.abs(a), axis=1) + 2)
    scaling = np.geomspace(0.01, 100, 18)
    a *= scaling[:, None] * scaling[None, :]
    for tolerance in [0.0, 0.003, 0.1]:
        P = oracle.fsai(csr(oracle, a, dtype, block=3), max_row_size=8, kap_tolerance=tolerance,
                        max_step_size=step)
        np.testing.assert_allclose(P["G"].to_dense(), greedy_reference(a, 8, step, tolerance),
                                   rtol=3e-5, atol=1e-6)


def test_greedy_search_on_random_sparse_spd(oracle):
    # Larger random systems, every width/step combination: the C restatement and the
    # numpy restatement must pick the same supports.
    rng = np.random.default_rng(5)
    for trial in range(6):
        n = 40
        a = ss.random(n, n, density=0.12, random_state=rng).toarray()
        a = a + a.T
        np.fill_diagonal(a, np.abs(a).sum(axis=1) + 0.5)
        for width, step, tol in [(6, 1, 0.0), (8, 2, 1e-3), (12, 3, 0.0), (5, 5, 0.05)]:
            P = oracle.fsai(csr(oracle, a), max_row_size=width, kap_tolerance=tol, max_step_size=step)
            expected = greedy_reference(a, width, step, tol)
            got = P["G"].to_dense()
            np.testing.assert_array_equal(got != 0, expected != 0)
            np.testing.assert_allclose(got, expected, rtol=1e-10, atol=1e-13)


def test_batch_width_cap_and_pivot_guard(oracle):
    a = np.array([[1.0, 1.0 - 1e-14, 0.0], [1.0 - 1e-14, 1.0, 0.0], [0.0, 0.0, 2.0]])
    P = oracle.fsai(csr(oracle, a), max_row_size=2, max_step_size=6)
    assert P["truncated_rows"] == 1
    np.testing.assert_allclose(P["G"].to_dense(), np.diag(1 / np.sqrt(np.diag(a))))
    P = oracle.fsai(csr(oracle, a), max_row_size=1, max_step_size=6)
    np.testing.assert_allclose(P["G"].to_dense(), np.diag(1 / np.sqrt(np.diag(a))))


@pytest.mark.parametrize("step", [2, 3])
def test_batch_grows_through_narrow_frontier(oracle, step):
    a = np.diag(np.full(12, 3.0)) + np.diag(-np.ones(11), 1) + np.diag(-np.ones(11), -1)
    P = oracle.fsai(csr(oracle, a), max_row_size=8, max_step_size=step, kap_tolerance=0.0)
    g = P["G"].to_dense()
    np.testing.assert_array_equal(np.count_nonzero(g, axis=1), np.minimum(np.arange(1, 13), 8))
    for i in range(12):
        support = np.arange(max(0, i - 7), i + 1)
        rhs = np.zeros(len(support))
        rhs[-1] = 1
        z = np.linalg.solve(a[np.ix_(support, support)], rhs)
        np.testing.assert_allclose(g[i, support], z / np.sqrt(z[-1]), rtol=1e-12)


def test_zero_tolerance_disables_roundoff_stopping(oracle):
    n = 32
    a = np.diag(np.full(n, 3.0)) + np.diag(-np.ones(n - 1), 1) + np.diag(-np.ones(n - 1), -1)
    P = oracle.fsai(csr(oracle, a), max_row_size=n, kap_tolerance=0, max_step_size=3)
    np.testing.assert_array_equal(np.diff(P["G"].off), np.arange(1, n + 1))
    g = P["G"].to_dense()
    np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=1e-12, atol=1e-15)


def test_structure_pack_and_transpose(oracle):
    rng = np.random.default_rng(3)
    a = ss.random(60, 60, density=0.1, random_state=rng).toarray()
    a = a + a.T
    np.fill_diagonal(a, np.abs(a).sum(axis=1) + 1)
    P = oracle.fsai(csr(oracle, a), max_row_size=7)
    G, GT = P["G"], P["GT"]
    for i in range(60):
        cols = G.col[G.off[i]:G.off[i + 1]]
        assert np.all(np.diff(cols) > 0) and cols[-1] == i  # sorted, diagonal last
        tcols = GT.col[GT.off[i]:GT.off[i + 1]]
        assert np.all(np.diff(tcols) > 0) and tcols[0] == i  # diagonal first
    np.testing.assert_array_equal(GT.to_dense(), G.to_dense().T)


def test_factor_dtype_conversion(oracle):
    # tests/test_squared_laplacian.py:92-109: float32 factor of a float64 matrix.
    with pytest.raises(ValueError, match="positive diagonal"):
        oracle.fsai(csr(oracle, np.eye(3) * 1e100), factor_dtype=np.float32)
    n = 25
    a = np.diag(np.full(n, 3.0)) + np.diag(np.full(n - 1, -1.0), 1) + np.diag(np.full(n - 1, -1.0), -1)
    P = oracle.fsai(csr(oracle, a), max_row_size=8, factor_dtype=np.float32)
    assert P["G"].dtype == np.float32
    np.testing.assert_array_equal(P["GT"].to_dense(), P["G"].to_dense().T)
    g = P["G"].to_dense()
    x = np.random.default_rng(72).normal(size=n)
    np.testing.assert_allclose(oracle.fsai_apply(P, x), g.T @ g @ x, rtol=1e-12, atol=1e-12)


def test_half_precision_factor(oracle):
    # tests/test_fsai_float16.py:19-33 and :38-56.
    rng = np.random.default_rng(19)
    r = rng.normal(size=(9, 9))
    a = (r @ r.T + 4 * np.eye(9)).astype(np.float16).astype(np.float64)
    P = oracle.fsai(csr(oracle, a, np.float16, block=3), max_row_size=9, kap_tolerance=0)
    assert P["G"].dtype == np.float16
    g = P["G"].to_dense()
    np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=1e-2, atol=1e-3)
    P = oracle.fsai(csr(oracle, np.diag([1, 0.01]), np.float16, block=2))
    y = oracle.fsai_apply(P, np.ones(2, np.float16))
    np.testing.assert_allclose(y.astype(np.float64), [1, 100], rtol=0, atol=0.01)
    for a in [np.array([[1, 2], [2, 4]]), np.array([[9, 21], [21, 49]])]:
        P = oracle.fsai(csr(oracle, a, np.float16), max_row_size=2)
        assert P["truncated_rows"] == 1
        np.testing.assert_allclose(P["G"].to_dense(), np.diag(1 / np.sqrt(np.diag(a))), rtol=2e-3)


def test_half_rounding_matches_numpy(oracle):
    # The binary16 emulation rounds exactly like numpy's float16 conversion.
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(size=2000) * 10.0 ** rng.integers(-9, 6, size=2000),
                        [6.1e-5, 5.96e-8, 2.98e-8, 65504.0, 65519.0, 65520.0, -0.0]])
    got = oracle.convert(x, np.float16)
    np.testing.assert_array_equal(got.view(np.uint16), x.astype(np.float16).view(np.uint16))
