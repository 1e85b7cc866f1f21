"""CG with the update step folded into the factored operator's second pass
(wpc_sqlap_cg_update): the curvature from the first pass's partials, the residual
update exactly the update step's arithmetic with that ALPHA, and whole solves
matching the separate-update path."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import NP, dev, host, laplacian_2d, upload

pytestmark = pytest.mark.gpu


def _vec(a, dtype):
    return torch.tensor(np.asarray(a), dtype=dtype, device=dev())


def _operator(side, seed, dtype, weight):
    import warp_preconditioners as wp
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(seed)
    k = laplacian_2d(side)
    n = k.shape[0]
    mass = rng.uniform(0.5, 1.5, size=n) / side ** 2
    A = wp.SquaredLaplacianOperator(upload(k, dtype), _vec(mass, dtype),
                                    torch.arange(n, dtype=torch.int32, device=dev()),
                                    row_lanes=1, mass_weight=weight)
    A._stream().fixed = StreamPlan.SLICED
    q = (k @ ss.diags(1 / mass) @ k + weight * ss.diags(mass)).tocsr()
    return n, A, q


def _state(nvec, inactive):
    from warp_preconditioners.linear import nat_state_fields

    f = nat_state_fields()
    st = np.zeros((f, nvec))
    st[0] = np.linspace(0.5, 2.0, nvec)  # RZ
    st[1] = 1.0                           # RR
    st[2] = 0.25                          # TARGET
    if inactive is not None:
        st[1, inactive] = 0.1             # RR <= TARGET: this slot rests
    return torch.tensor(st.reshape(-1), dtype=torch.float64, device=dev())


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("weight", [0.0, 1.0])
@pytest.mark.parametrize("nvec", [1, 3, 8])
def test_fused_step_is_the_update_step_with_the_energy_curvature(dtype, weight, nvec):
    from warp_preconditioners import _native as nat

    n, A, q = _operator(37, 5 + nvec, dtype, weight)
    rng = np.random.default_rng(nvec)
    p = _vec(rng.normal(size=n * nvec), dtype)
    r = _vec(rng.normal(size=n * nvec), dtype)
    inactive = 1 if nvec > 1 else None
    s = nat.stream()
    code = nat.code(dtype)

    # the separate path: A p with p.Ap partials, then the update step (x left alone)
    ap = torch.empty_like(p)
    part, nparts = A.apply_batch(p, ap, ap, 1.0, 0.0, nvec, dot_w=p)
    r_sep, st_sep = r.clone(), _state(nvec, inactive)
    rr_sep = torch.zeros(nat.lib().wpc_krylov_partials(n) * nvec, dtype=torch.float64,
                         device=dev())
    nat.call("wpc_cg_update_fold", n, nvec, None, nat.ptr(r_sep), nat.ptr(p), nat.ptr(ap), code,
             nparts, nat.ptr(part), nat.ptr(st_sep), nat.ptr(rr_sep), s)

    count = A.cg_update_parts(nvec, dtype)
    assert count == (n + 511) // 512
    r_f, st_f = r.clone(), _state(nvec, inactive)
    rr_f = torch.zeros(count * nvec, dtype=torch.float64, device=dev())
    A.cg_update(p, r_f, nvec, st_f, rr_f)
    torch.cuda.synchronize()

    F = st_f.numel() // nvec
    sf = host(st_f).reshape(F, nvec)
    ss_ = host(st_sep).reshape(F, nvec)
    # the curvature against the assembled operator on the host (float64)
    ph = host(p).astype(np.float64).reshape(n, nvec)
    exact = np.einsum("iv,iv->v", ph, q @ ph)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    np.testing.assert_allclose(sf[5], exact, rtol=tol)   # DENOM
    np.testing.assert_allclose(sf[5], ss_[5], rtol=tol)
    for f in (6, 7):                                      # RZ_OLD, ACTIVE
        np.testing.assert_array_equal(sf[f], ss_[f])
    active = sf[7] != 0
    np.testing.assert_array_equal(sf[3], np.where(active, sf[0] / sf[5], 0.0))  # ALPHA
    # r - ALPHA Ap in the vector type, Ap as the separate pass stores it
    T = NP[dtype]
    rh, aph = host(r).reshape(n, nvec), host(ap).reshape(n, nvec)
    want = rh.copy()
    for v in range(nvec):
        a = T(sf[3, v])
        if a != 0:
            want[:, v] = rh[:, v] - a * aph[:, v]
    got = host(r_f).reshape(n, nvec)
    np.testing.assert_array_equal(got, want)
    if inactive is not None:
        np.testing.assert_array_equal(got[:, inactive], rh[:, inactive])
    # r.r partials of the new residual
    rr = host(rr_f).reshape(count, nvec).sum(axis=0)
    np.testing.assert_allclose(rr, (got.astype(np.float64) ** 2).sum(axis=0), rtol=1e-12)


@pytest.mark.parametrize("weight", [0.0, 1.0])
def test_fused_solves_match_the_separate_update(weight):
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    n, A, q = _operator(48, 11, torch.float64, weight)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=torch.float32)
    rng = np.random.default_rng(3)
    B = _vec(rng.normal(size=(n, 3)), torch.float64)
    out = {}
    for fuse in (False, True):
        work = linear.KrylovWork("cg", A, P, n, 3, torch.float64, dev())
        work.fuse_update = fuse
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-10, maxiter=20000, check_every=25,
                                 work=work)
        assert (work._rr_parts is not None) == fuse  # the fused step ran
        out[fuse] = (o, host(X))
    xs = np.linalg.solve(q.toarray(), host(B))
    for fuse, (o, X) in out.items():
        for v in range(3):
            res = np.linalg.norm(host(B)[:, v] - q @ X[:, v]) / np.linalg.norm(host(B)[:, v])
            assert res <= 1.5e-10, (fuse, v, res)
    its_f, its_s = out[True][0]["iterations"], out[False][0]["iterations"]
    for a, b in zip(its_f, its_s):
        assert abs(a - b) <= max(2, 0.02 * b), (its_f, its_s)
    for v in range(3):
        e = np.linalg.norm(out[True][1][:, v] - xs[:, v]) / np.linalg.norm(xs[:, v])
        assert e <= 1e-6


def test_fused_steps_replay_from_a_graph_as_run_eagerly():
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    n, A, q = _operator(40, 2, torch.float64, 1.0)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=torch.float32)
    B = _vec(np.random.default_rng(9).normal(size=(n, 2)), torch.float64)
    xs = []
    for graph in (False, True):
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-9, maxiter=5000, check_every=25,
                                 use_cuda_graph=graph)
        assert o["work"]._rr_parts is not None
        xs.append((host(X), o["iterations"]))
    np.testing.assert_array_equal(xs[0][0], xs[1][0])
    assert xs[0][1] == xs[1][1]


@pytest.mark.parametrize("weight", [0.0, 1.0])
def test_fused_solves_with_scalar_folds(weight, monkeypatch):
    """Partial counts too large for the in-place folds (forced here): the curvature
    folded by wpc_cg_alpha, the second pass taking ALPHA from the state and
    advancing x, then wpc_cg_beta and wpc_cg_direction -- the same solves as the
    separate update step with the same scalar folds."""
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    monkeypatch.setattr(linear, "FOLD_IN_PLACE", 1)
    n, A, q = _operator(44, 21, torch.float64, weight)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=torch.float32)
    rng = np.random.default_rng(7)
    B = _vec(rng.normal(size=(n, 3)), torch.float64)
    out = {}
    for fuse in (False, True):
        work = linear.KrylovWork("cg", A, P, n, 3, torch.float64, dev())
        work.fuse_update = fuse
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-10, maxiter=20000, check_every=25,
                                 work=work)
        assert (work._rr_parts is not None) == fuse
        out[fuse] = (o["iterations"], host(X))
    for fuse, (its, X) in out.items():
        for v in range(3):
            res = np.linalg.norm(host(B)[:, v] - q @ X[:, v]) / np.linalg.norm(host(B)[:, v])
            assert res <= 1.5e-10, (fuse, v, res)
    for a, b in zip(out[True][0], out[False][0]):
        assert abs(a - b) <= max(2, 0.02 * b), (out[True][0], out[False][0])
    # replayed from a graph as run eagerly
    xs = []
    for graph in (False, True):
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-9, maxiter=5000, check_every=25,
                                 use_cuda_graph=graph)
        xs.append((host(X), o["iterations"]))
    np.testing.assert_array_equal(xs[0][0], xs[1][0])
    assert xs[0][1] == xs[1][1]


def test_energy_partials_never_move_for_up_to_8_vectors():
    """Graphs captured by served requests bake in this buffer: it is sized for 8
    vectors when first made, so a request with more vectors never reallocates it
    under a graph another request has in flight."""
    from warp_preconditioners import _native as nat

    n, A, _ = _operator(30, 1, torch.float64, 1.0)
    gen = nat.generation()
    first = A.energy_partials(1)
    assert first.numel() >= nat.lib().wpc_sqlap_cg_partials(n) * 8
    after = nat.generation()
    for nvec in range(1, 9):
        assert A.energy_partials(nvec).data_ptr() == first.data_ptr()
    assert nat.generation() == after and after >= gen


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("weight", [0.0, 1.0])
def test_second_pass_advances_x(dtype, weight):
    """With x the second pass also applies x += ALPHA p (in the vector type, only
    for stepping slots), leaving r and the state exactly as without x."""
    nvec = 3
    n, A, _ = _operator(29, 4, dtype, weight)
    rng = np.random.default_rng(12)
    p = _vec(rng.normal(size=n * nvec), dtype)
    r = _vec(rng.normal(size=n * nvec), dtype)
    x0 = _vec(rng.normal(size=n * nvec), dtype)
    count = A.cg_update_parts(nvec, dtype)
    outs = []
    for with_x in (False, True):
        rr = torch.zeros(count * nvec, dtype=torch.float64, device=dev())
        st, rv, xv = _state(nvec, 1), r.clone(), x0.clone()
        A.cg_update(p, rv, nvec, st, rr, x=xv if with_x else None)
        torch.cuda.synchronize()
        outs.append((host(rv), host(st), host(xv), host(rr)))
    for a, b in zip(outs[0][:2], outs[1][:2]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(outs[0][3], outs[1][3])
    np.testing.assert_array_equal(outs[0][2], host(x0))  # untouched without x
    F = outs[1][1].size // nvec
    alpha = outs[1][1].reshape(F, nvec)[3]
    T = NP[dtype]
    want = host(x0).reshape(n, nvec).copy()
    ph = host(p).reshape(n, nvec)
    for v in range(nvec):
        a = T(alpha[v])
        if a != 0:
            want[:, v] = want[:, v] + a * ph[:, v]
    np.testing.assert_array_equal(outs[1][2].reshape(n, nvec), want)
    assert alpha[1] == 0.0  # the resting slot's x is unchanged
