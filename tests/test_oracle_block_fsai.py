"""The oracle's block FSAI (benchmarks/block_fsai_candidate.py) pinned by the
reference's own dense checks (tests/test_fsai_candidates.py:20-55,
tests/test_block_challenge.py:65-90): every block row equals the normalized
principal inverse of its selected support, G A G^T = I per block row, and the full
block pattern reproduces A^-1."""

import numpy as np
import pytest
import scipy.sparse as ss


def _dense(C):
    return ss.csr_matrix((C.val, C.col, C.off), shape=(C.nrow, C.ncol)).toarray()


def _spd(rng, n, scaled=True):
    r = rng.normal(size=(n, n))
    a = r @ r.T + np.eye(n)
    if scaled:
        s = np.geomspace(0.01, 100, n)
        a *= s[:, None] * s[None, :]
    return a


def _check_rows(g, a, bs, nblocks, rtol=2e-11, atol=1e-10):
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
def test_block_principal_inverse(oracle, bs, blocks):
    rng = np.random.default_rng(712)
    n = 4 * bs
    a = _spd(rng, n)
    P = oracle.block_fsai(oracle.CSR.from_scipy(a, np.float64, bs), max_blocks=blocks)
    g = _dense(P["G"])
    assert np.linalg.eigvalsh(g.T @ g).min() > 0
    _check_rows(g, a, bs, 4)
    if blocks == 4:
        np.testing.assert_allclose(g.T @ g, np.linalg.inv(a), rtol=2e-11, atol=1e-10)
    if blocks == 1:
        for i in range(4):
            rows = slice(bs * i, bs * (i + 1))
            np.testing.assert_allclose((g.T @ g)[rows, rows], np.linalg.inv(a[rows, rows]),
                                       rtol=2e-11)


def test_block_refit_selected_principal_inverse(oracle):
    rng = np.random.default_rng(992)
    old = _spd(rng, 12, scaled=False) + 2 * np.eye(12)
    new = _spd(rng, 12, scaled=False) + 4 * np.eye(12)
    P = oracle.block_fsai(oracle.CSR.from_scipy(old, np.float64, 3), max_blocks=2)
    Q = oracle.block_fsai(oracle.CSR.from_scipy(new, np.float64, 3), max_blocks=2, refit_from=P)
    np.testing.assert_array_equal(P["G"].col, Q["G"].col)
    _check_rows(_dense(Q["G"]), new, 3, 4, rtol=2e-12, atol=2e-13)


def test_block_greedy_selection_matches_brute_force(oracle):
    # The block added at each step maximizes the block residual score over the
    # strictly lower block frontier (ties: smaller block).  Brute force on dense
    # blocks with an explicit sparsity pattern.
    rng = np.random.default_rng(5)
    bs, nb = 2, 7
    n = bs * nb
    a = _spd(rng, n, scaled=False) + 3 * np.eye(n)
    mask = rng.random((nb, nb)) < 0.5
    mask = mask | mask.T
    np.fill_diagonal(mask, True)
    a = a * np.kron(mask, np.ones((bs, bs)))
    a = a + n * np.eye(n)  # keep it SPD after masking
    P = oracle.block_fsai(oracle.CSR.from_scipy(a, np.float64, bs), max_blocks=2)
    g = _dense(P["G"])
    d = 1 / np.sqrt(np.diag(a))
    at = a * d[:, None] * d[None, :]
    for i in range(nb):
        own = np.arange(bs * i, bs * (i + 1))
        z = np.linalg.solve(at[np.ix_(own, own)], np.eye(bs))
        frontier = sorted({j for j in range(i) if mask[i, j]})
        scores = [np.sum((at[np.ix_(np.arange(bs * j, bs * j + bs), own)] @ z) ** 2)
                  for j in frontier]
        chosen = sorted({c // bs for c in np.flatnonzero(np.any(g[own] != 0, axis=0))} - {i})
        if not frontier or max(scores) == 0:
            assert chosen == []
            continue
        best = max(scores)
        want = min(j for j, s in zip(frontier, scores) if s == best)
        assert chosen == [want]
