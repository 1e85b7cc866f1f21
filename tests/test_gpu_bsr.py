"""GPU parity of the 3x3 sliced block product (wpc_bsr3_sell_stream): bit for bit
against the oracle's bsr_spmv (block product, then accumulation, in the vector
type), with fused dot partials, NaN/inf rules, and CG through the block operator."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host

pytestmark = pytest.mark.gpu


def _pattern(rng, nb, lengths):
    """Block pattern whose row i holds lengths[i] distinct columns (sorted) within a
    band of +-400 block columns."""
    rows, cols = [], []
    for i, m in enumerate(lengths):
        lo, hi = max(0, i - 400), min(nb, i + 400)
        c = np.sort(rng.choice(np.arange(lo, hi), size=min(m, hi - lo), replace=False))
        rows += [i] * len(c)
        cols += c.tolist()
    p = ss.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(nb, nb))
    p.sort_indices()
    return p


def _upload(p, data, vdtype):
    from warp_preconditioners.sparse import BsrMatrix

    nb = p.shape[0]
    return BsrMatrix(nb, nb, (3, 3), torch.tensor(p.indptr.astype(np.int32), device=dev()),
                     torch.tensor(p.indices.astype(np.int32), device=dev()),
                     torch.tensor(data, dtype=vdtype, device=dev()))


def _oracle(oracle, p, data):
    return oracle.CSR(p.indptr.astype(np.int32), p.indices.astype(np.int32),
                      np.ascontiguousarray(data), p.shape[0], p.shape[0], 3)


@pytest.mark.parametrize("col16", [True, False])
@pytest.mark.parametrize("types", ["f64/f64", "f32/f64", "f32/f32"])
@pytest.mark.parametrize("nvec", [1, 2, 3, 4])
def test_bsr3_sell_matches_oracle_bitwise(oracle, monkeypatch, col16, types, nvec):
    # Block rows of every length 0..17 (every full and masked group of the round
    # trips), empty rows, alpha and beta != 0, against the oracle's serial block
    # product bit for bit, per right-hand side.
    from warp_preconditioners.sparse import BsrPlan

    monkeypatch.setattr(BsrPlan, "COL16", col16)
    vt, xt = types.split("/")
    npv = np.float64 if vt == "f64" else np.float32
    npx = np.float64 if xt == "f64" else np.float32
    rng = np.random.default_rng(7 + nvec)
    nb = 1000
    lengths = [(i * 7) % 18 for i in range(nb)]
    p = _pattern(rng, nb, lengths)
    data = rng.normal(size=(p.nnz, 3, 3)).astype(npv)
    A = _upload(p, data, torch.float64 if vt == "f64" else torch.float32)
    plan = BsrPlan(A)
    assert plan.col16 == col16
    X = rng.normal(size=(nb * 3, nvec)).astype(npx)
    Y = rng.normal(size=(nb * 3, nvec)).astype(npx)
    td = torch.float64 if xt == "f64" else torch.float32
    Z = torch.empty(nb * 3 * nvec, dtype=td, device=dev())
    W = torch.tensor(rng.normal(size=nb * 3 * nvec).astype(npx), device=dev())
    out = torch.zeros(nvec, dtype=torch.float64, device=dev())
    plan.run(torch.tensor(X.ravel(), device=dev()), torch.tensor(Y.ravel(), device=dev()), Z,
             0.7, -0.3, nvec, w=W, out=out)
    got = host(Z).reshape(nb * 3, nvec)
    Ao = _oracle(oracle, p, data)
    for v in range(nvec):
        ref = oracle.bsr_spmv(Ao, np.ascontiguousarray(X[:, v]), np.ascontiguousarray(Y[:, v]),
                              0.7, -0.3, vec_dtype=npx)
        np.testing.assert_array_equal(got[:, v], ref)
    dots = (host(W).reshape(nb * 3, nvec).astype(np.float64) * got.astype(np.float64)).sum(0)
    np.testing.assert_allclose(host(out), dots, rtol=1e-12)


def test_bsr3_sell_nan_inf_rules(oracle):
    # alpha = 0 never reads x, beta = 0 never reads y, and padding adds nothing even
    # where x holds inf next to a short row (0 * inf would be NaN).
    from warp_preconditioners.sparse import BsrPlan

    rng = np.random.default_rng(3)
    nb = 200
    lengths = [0 if i == 7 else (1 if i == 70 else 9) for i in range(nb)]
    p = _pattern(rng, nb, lengths).tolil()
    p[70, :] = 0
    p[70, 71] = 1
    p = p.tocsr()
    p.eliminate_zeros()
    p.sort_indices()
    data = rng.normal(size=(p.nnz, 3, 3))
    A = _upload(p, data, torch.float64)
    plan = BsrPlan(A)
    x = rng.normal(size=nb * 3)
    x[71 * 3] = np.inf  # one component: each block product then holds one inf term
    x[0:3] = np.nan
    y = rng.normal(size=nb * 3)
    z = torch.empty(nb * 3, dtype=torch.float64, device=dev())
    plan.run(torch.tensor(x, device=dev()), torch.tensor(y, device=dev()), z, 1.5, 0.5)
    got = host(z)
    ref = oracle.bsr_spmv(_oracle(oracle, p, data), x, y, 1.5, 0.5)
    np.testing.assert_array_equal(got, ref)
    assert np.array_equal(got[21:24], 0.5 * y[21:24])  # the empty block row 7
    assert np.isinf(got[210:213]).all()                 # block row 70: inf, not NaN
    xn = torch.full((nb * 3,), float("nan"), dtype=torch.float64, device=dev())
    plan.run(xn, torch.tensor(y, device=dev()), z, 0.0, 2.0)
    np.testing.assert_array_equal(host(z), 2.0 * y)
    yn = torch.full((nb * 3,), float("nan"), dtype=torch.float64, device=dev())
    plan.run(torch.tensor(np.nan_to_num(x, nan=0.0, posinf=0.0), device=dev()), yn, z, 1.0, 0.0)
    assert np.isfinite(host(z)).all()


def test_block_operator_cg_against_oracle(oracle):
    # CG with FSAI on a K (x) C block system, the operator running the 3x3 block
    # product (fused dot partials in place): iterations and iterate follow the
    # oracle's PCG on the same (scalarised) matrix; the block operator's choice is
    # recorded with both timings.
    import warp_preconditioners as wp
    from warp_preconditioners import linear, workloads
    from warp_preconditioners.sparse import BsrMatrix

    off, col, val, _, _ = workloads.kron_coupling(12, 1.0e3, seed=0)
    nb = len(off) - 1
    A = BsrMatrix(nb, nb, (3, 3), torch.tensor(off, device=dev()), torch.tensor(col, device=dev()),
                  torch.tensor(val, device=dev()))
    op = linear.aslinearoperator(A)
    assert op.fused
    b = np.random.default_rng(0).standard_normal(nb * 3)
    P = wp.FSAI(A, max_row_size=8)
    B = torch.tensor(b[:, None], device=dev())
    X = torch.zeros_like(B)
    out = linear.solve_batched("cg", op, B, X, M=P, tol=1e-10, maxiter=5000, check_every=1)
    assert out["work"].use_fused
    assert op.block_mode(1, torch.float64) in ("bsr", "csr")
    assert (1, torch.float64) in op.block_timings or op.block_mode(1, torch.float64) == "bsr"
    a = ss.bsr_matrix((val, col, off), shape=(nb * 3, nb * 3)).tocsr()
    Ao = oracle.CSR.from_scipy(a, np.float64)
    Po = oracle.fsai(Ao, max_row_size=8)
    xo, ito, _ = oracle.krylov("cg", b, A=Ao, P=Po, tol=1e-10, maxiter=5000, check_every=1)
    assert abs(out["iterations"][0] - ito) <= max(2, 0.02 * ito)
    np.testing.assert_allclose(host(X[:, 0]), xo, rtol=1e-7, atol=1e-9 * np.abs(xo).max())


@pytest.mark.parametrize("size", [1, 2, 3, 4])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_generic_bsr_spmv_against_oracle(oracle, size, dtype):
    # wpc_bsr_spmv (b x b blocks, 8 lanes per block row and a lane reduction: a
    # different summation order) against the oracle's product within rounding,
    # exactly on small-integer data, with empty block rows and the alpha/beta rules.
    from warp_preconditioners import _native as nat
    from warp_preconditioners.sparse import BsrMatrix

    npd = np.float64 if dtype == torch.float64 else np.float32
    rng = np.random.default_rng(200 + size)
    nb = 300
    p = _pattern(rng, nb, [0 if i % 37 == 5 else 1 + (i % 11) for i in range(nb)])
    for kind in ("integer", "random"):
        if kind == "integer":
            data = rng.integers(-4, 5, size=(p.nnz, size, size)).astype(npd)
            x, y = rng.integers(-3, 4, size=(2, nb * size)).astype(npd)
        else:
            data = rng.normal(size=(p.nnz, size, size)).astype(npd)
            x, y = rng.normal(size=(2, nb * size)).astype(npd)
        A = BsrMatrix(nb, nb, (size, size), torch.tensor(p.indptr.astype(np.int32), device=dev()),
                      torch.tensor(p.indices.astype(np.int32), device=dev()),
                      torch.tensor(data, device=dev()))
        z = torch.empty(nb * size, dtype=dtype, device=dev())
        xd, yd = torch.tensor(x, device=dev()), torch.tensor(y, device=dev())  # (kept alive)
        nat.call("wpc_bsr_spmv", nb, size, nat.ptr(A.offsets), nat.ptr(A.columns),
                 nat.ptr(A.values), nat.code(dtype), nat.ptr(xd), nat.ptr(yd), nat.ptr(z), 2.0,
                 -0.5, nat.stream())
        Ao = oracle.CSR(p.indptr.astype(np.int32), p.indices.astype(np.int32),
                        np.ascontiguousarray(data), nb, nb, size)
        ref = oracle.bsr_spmv(Ao, x, y, 2.0, -0.5)
        if kind == "integer":
            np.testing.assert_array_equal(host(z), ref)
        else:
            tol = 2e-5 if dtype == torch.float32 else 1e-13
            np.testing.assert_allclose(host(z), ref, rtol=tol, atol=tol)
