"""GPU parity of the streamed CSR product and the fused CG path."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host, oracle_csr, random_spd, upload

pytestmark = pytest.mark.gpu


def _plan(A):
    from warp_preconditioners.sparse import StreamPlan

    return StreamPlan(A)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_stream_matches_oracle_bitwise(oracle, dtype):
    # Serial row sums in storage order without fused multiply-add: bit-identical
    # to the oracle's serial product, including an empty row and a long row.
    rng = np.random.default_rng(4)
    a = ss.random(3000, 3000, density=0.004, random_state=rng, format="lil")
    a[0, :] = 0
    a[17, :] = rng.normal(size=3000)  # 3000 entries: a block of its own
    a = a.tocsr()
    npd = np.float32 if dtype == torch.float32 else np.float64
    A = upload(a, dtype)
    plan = _plan(A)
    x = rng.normal(size=3000).astype(npd)
    y = rng.normal(size=3000).astype(npd)
    z = torch.empty(3000, dtype=dtype, device=dev())
    plan.run(torch.tensor(x, device=dev()), torch.tensor(y, device=dev()), z, 0.7, -0.3)
    ref = oracle.spmv(oracle_csr(oracle, a, npd), x, y, 0.7, -0.3)
    got = host(z)
    normal = np.ones(3000, bool)
    normal[17] = False  # the long row is reduced by a tree, not serially
    np.testing.assert_array_equal(got[normal], ref[normal])
    np.testing.assert_allclose(got[17], ref[17], rtol=1e-5 if dtype == torch.float32 else 1e-12)


@pytest.mark.parametrize("nvec", [1, 2, 3, 4, 5, 6, 7, 8])
def test_stream_batched_dot_and_determinism(nvec):
    rng = np.random.default_rng(nvec)
    a = random_spd(rng, 5000, 0.002)
    A = upload(a)
    plan = _plan(A)
    X = torch.tensor(rng.normal(size=(5000, nvec)), device=dev())
    W = torch.tensor(rng.normal(size=(5000, nvec)), device=dev())
    Z = torch.empty_like(X)
    out = torch.zeros(nvec, dtype=torch.float64, device=dev())
    plan.run(X.reshape(-1), None, Z.reshape(-1), 1.0, 0.0, nvec, w=W.reshape(-1), out=out)
    expected = a @ host(X)
    np.testing.assert_allclose(host(Z), expected, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(host(out), (host(W) * expected).sum(0), rtol=1e-10)
    first = host(out).copy()
    for _ in range(3):
        plan.run(X.reshape(-1), None, Z.reshape(-1), 1.0, 0.0, nvec, w=W.reshape(-1), out=out)
        np.testing.assert_array_equal(host(out), first)


def test_fused_and_unfused_cg_agree():
    from warp_preconditioners import FSAI, linear

    rng = np.random.default_rng(2)
    a = random_spd(rng, 4000, 0.002, (1e-2, 1e2))
    A = upload(a)
    P = FSAI(A, max_row_size=8)
    B = torch.tensor(rng.normal(size=(4000, 3)), device=dev())
    X1 = torch.zeros_like(B)
    out1 = linear.solve_batched("cg", A, B, X1, M=P, tol=1e-10, check_every=1)
    assert out1["work"].use_fused
    X2 = torch.zeros_like(B)
    Pu = FSAI(A, max_row_size=8, apply_lanes=8)
    Pu.fused = False
    out2 = linear.solve_batched("cg", A, B, X2, M=Pu, tol=1e-10, check_every=1)
    assert not out2["work"].use_fused
    for i1, i2 in zip(out1["iterations"], out2["iterations"]):
        assert abs(i1 - i2) <= 1
    np.testing.assert_allclose(host(X1), host(X2), rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(a @ host(X1), host(B), rtol=0, atol=1e-9 * np.abs(host(B)).max())


def test_batched_equals_single_rhs_solves():
    # Each right-hand side converges to the same iterate alone and in the batch.
    from warp_preconditioners import FSAI, linear

    rng = np.random.default_rng(12)
    a = random_spd(rng, 3000, 0.003)
    A = upload(a)
    P = FSAI(A, max_row_size=8)
    B = torch.tensor(rng.normal(size=(3000, 4)), device=dev())
    X = torch.zeros_like(B)
    outb = linear.solve_batched("cg", A, B, X, M=P, tol=1e-9, check_every=1)
    for v in range(4):
        b = B[:, v].contiguous()
        x = torch.zeros_like(b)
        out = linear.solve_batched("cg", A, b[:, None], x[:, None], M=P, tol=1e-9, check_every=1)
        assert out["iterations"][0] == outb["iterations"][v]
        np.testing.assert_allclose(host(x), host(X[:, v]), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("nvec", [5, 6, 7])
def test_batched_cg_odd_vector_counts(nvec):
    from warp_preconditioners import FSAI, linear

    rng = np.random.default_rng(10 + nvec)
    a = random_spd(rng, 3000, 0.003)
    A = upload(a)
    P = FSAI(A, max_row_size=8)
    B = torch.tensor(rng.normal(size=(3000, nvec)), device=dev())
    X = torch.zeros_like(B)
    out = linear.solve_batched("cg", A, B, X, M=P, tol=1e-10, check_every=5)
    assert out["work"].use_fused
    np.testing.assert_allclose(host(X), np.linalg.solve(a, host(B)), rtol=1e-7, atol=1e-9)
