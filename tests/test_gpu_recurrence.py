"""CG in recurrence form (wpc_sqlap_cg_direction_energy): the direction step of
one step folded into the next step's first pass of the factored operator, which
gathers z and carries t = M^-1 L p across steps.  The kernel against the direction
step followed by the plain first pass, and whole solves (eager, replayed from
graphs, float32) against the five-launch sequence."""

import numpy as np
import pytest
import torch

from gpu_helpers import NP, dev, host, upload
from test_gpu_fused_update import _operator, _state, _vec

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("weight", [0.0, 1.0])
@pytest.mark.parametrize("nvec", [1, 3, 8])
def test_direction_energy_is_the_direction_step_then_the_first_pass(dtype, weight, nvec):
    from warp_preconditioners import _native as nat

    n, A, q = _operator(37, 30 + nvec, dtype, weight)
    rng = np.random.default_rng(40 + nvec)
    p = _vec(rng.normal(size=n * nvec), dtype)
    z = _vec(rng.normal(size=n * nvec), dtype)
    count = A.cg_update_parts(nvec, dtype)
    m_parts = 700
    part_rz = _vec(rng.uniform(0.5, 1.5, size=m_parts * nvec) / m_parts, torch.float64)
    part_rr = _vec(rng.uniform(0.5, 1.5, size=count * nvec) / count, torch.float64)
    inactive = 1 if nvec > 1 else None
    base = _state(nvec, inactive)
    F = base.numel() // nvec
    b = host(base).reshape(F, nvec)
    b[4] = 5.0                                   # ITER
    b[6] = np.linspace(1.0, 3.0, nvec)           # RZ_OLD
    b[7] = 1.0                                   # ACTIVE
    if inactive is not None:
        b[7, inactive] = 0.0
    base = torch.tensor(b.reshape(-1), dtype=torch.float64, device=dev())
    s = nat.stream()
    code = nat.code(dtype)
    scratch_r = torch.zeros_like(p)
    rr_tmp = torch.zeros(count * nvec, dtype=torch.float64, device=dev())

    def first_pass(pv, t):
        A.cg_update(pv, scratch_r, nvec, base.clone(), rr_tmp, passes=1, t=t)
        return A.energy_partials(nvec)[:count * nvec].clone()

    t0 = torch.empty_like(p)
    first_pass(p, t0)
    # the direction step, then the plain first pass on its p
    p_ref, st_ref, t_ref = p.clone(), base.clone(), torch.empty_like(p)
    nat.call("wpc_cg_direction_fold", n, nvec, nat.ptr(z), nat.ptr(p_ref), None, code, m_parts,
             nat.ptr(part_rz), count, nat.ptr(part_rr), nat.ptr(st_ref), s)
    e_ref = first_pass(p_ref, t_ref)
    # the recurrence pass
    p_rec, st_rec, t_rec = p.clone(), base.clone(), t0.clone()
    A.direction_energy(z, p_rec, t_rec, nvec, m_parts, part_rz, part_rr, st_rec)
    e_rec = A.energy_partials(nvec)[:count * nvec].clone()
    torch.cuda.synchronize()

    sr, sf = host(st_ref).reshape(F, nvec), host(st_rec).reshape(F, nvec)
    for f in (0, 1, 8, 9, 10):                   # RZ RR RR_NEW RZ_NEW BETA (fold order)
        np.testing.assert_allclose(sf[f], sr[f], rtol=1e-14, atol=0)
    for f in (2, 3, 4, 5, 6, 7, 11):
        np.testing.assert_array_equal(sf[f], sr[f])
    fz = host(part_rz).reshape(m_parts, nvec).sum(axis=0)
    np.testing.assert_allclose(sf[9], fz, rtol=1e-13)
    T = NP[dtype]
    ph, zh = host(p).reshape(n, nvec), host(z).reshape(n, nvec)
    pr, tr = host(p_rec).reshape(n, nvec), host(t_rec).reshape(n, nvec)
    t0h = host(t0).reshape(n, nvec)
    tref = host(t_ref).reshape(n, nvec)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    for v in range(nvec):
        if sf[7, v] == 0:                         # resting: p and t kept
            np.testing.assert_array_equal(pr[:, v], ph[:, v])
            np.testing.assert_array_equal(tr[:, v], t0h[:, v])
            assert sf[4, v] == 5.0
            continue
        assert sf[4, v] == 6.0
        beta = T(sf[10, v])
        np.testing.assert_array_equal(pr[:, v], zh[:, v] + beta * ph[:, v])
        # t = M^-1 L p by the recurrence, against the plain pass on the same p
        err = np.linalg.norm(tr[:, v].astype(np.float64) - tref[:, v]) / np.linalg.norm(tref[:, v])
        assert err <= tol, (v, err)
    # the curvature: partial sums against the plain pass and the assembled operator
    er = host(e_rec).reshape(count, nvec).sum(axis=0)
    ef = host(e_ref).reshape(count, nvec).sum(axis=0)
    np.testing.assert_allclose(er, ef, rtol=tol)
    p64 = pr.astype(np.float64)
    exact = np.einsum("iv,iv->v", p64, q @ p64)
    np.testing.assert_allclose(er, exact, rtol=tol)


def _counted(A):
    calls = []
    inner = A.direction_energy

    def wrapped(*args):
        calls.append(1)
        return inner(*args)

    A.direction_energy = wrapped
    return calls


@pytest.mark.parametrize("weight", [0.0, 1.0])
def test_recurrence_solves_match_the_direction_step(weight):
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    n, A, q = _operator(48, 61, torch.float64, weight)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=torch.float32)
    B = _vec(np.random.default_rng(5).normal(size=(n, 3)), torch.float64)
    calls = _counted(A)
    out = {}
    for rec in (False, True):
        work = linear.KrylovWork("cg", A, P, n, 3, torch.float64, dev())
        work.recurrence = rec
        before = len(calls)
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-10, maxiter=20000, check_every=25,
                                 work=work)
        assert (len(calls) > before) == rec
        out[rec] = (o["iterations"], host(X))
    Bh = host(B)
    for rec, (its, X) in out.items():
        for v in range(3):
            res = np.linalg.norm(Bh[:, v] - q @ X[:, v]) / np.linalg.norm(Bh[:, v])
            assert res <= 1.5e-10, (rec, v, res)
    for a, b in zip(out[True][0], out[False][0]):
        assert abs(a - b) <= max(2, 0.02 * b), (out[True][0], out[False][0])


def test_recurrence_graph_replays_match_eager_steps():
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    n, A, q = _operator(40, 62, torch.float64, 1.0)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=torch.float32)
    B = _vec(np.random.default_rng(19).normal(size=(n, 2)), torch.float64)
    calls = _counted(A)
    xs = []
    for graph in (False, True):
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-9, maxiter=5000, check_every=25,
                                 use_cuda_graph=graph)
        torch.cuda.synchronize()
        xs.append((host(X), o["iterations"], host(o["work"].state)))
    assert calls
    np.testing.assert_array_equal(xs[0][0], xs[1][0])
    assert xs[0][1] == xs[1][1]
    np.testing.assert_array_equal(xs[0][2], xs[1][2])


def test_recurrence_float32_solve():
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    n, A, q = _operator(40, 63, torch.float32, 1.0)
    P = wp.FSAI(upload(q, torch.float32), max_row_size=8, factor_dtype=torch.float32)
    B = _vec(np.random.default_rng(23).normal(size=(n, 3)), torch.float32)
    calls = _counted(A)
    out = {}
    for rec in (False, True):
        work = linear.KrylovWork("cg", A, P, n, 3, torch.float32, dev())
        work.recurrence = rec
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-5, maxiter=20000, check_every=25,
                                 work=work)
        out[rec] = (o["iterations"], host(X).astype(np.float64))
    assert calls
    # float32 on the squared Laplacian stalls near kappa * eps in the true residual
    # (6e-4 here for either form): the recurrence must do as well as the plain form
    Bh = host(B).astype(np.float64)
    res = {rec: [np.linalg.norm(Bh[:, v] - q @ X[:, v]) / np.linalg.norm(Bh[:, v])
                 for v in range(3)] for rec, (_, X) in out.items()}
    for v in range(3):
        assert res[False][v] <= 5e-3, res
        assert res[True][v] <= 2.0 * res[False][v] + 1e-5, res
    for a, b in zip(out[True][0], out[False][0]):
        assert abs(a - b) <= max(3, 0.05 * b), (out[True][0], out[False][0])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("nvec", [1, 3, 8])
def test_direction_energy_with_beta_from_the_state(dtype, nvec):
    """part_rz None (partial counts too large to fold in every workgroup): BETA and
    ACTIVE as wpc_cg_beta left them, p exactly as wpc_cg_direction writes it, the
    state untouched."""
    from warp_preconditioners import _native as nat

    n, A, q = _operator(33, 50 + nvec, dtype, 1.0)
    rng = np.random.default_rng(60 + nvec)
    p = _vec(rng.normal(size=n * nvec), dtype)
    z = _vec(rng.normal(size=n * nvec), dtype)
    count = A.cg_update_parts(nvec, dtype)
    inactive = 1 if nvec > 1 else None
    base = _state(nvec, inactive)
    F = base.numel() // nvec
    b = host(base).reshape(F, nvec)
    b[7] = 1.0                                    # ACTIVE
    b[10] = np.linspace(0.3, 0.9, nvec)           # BETA
    if inactive is not None:
        b[7, inactive], b[10, inactive] = 0.0, 0.0
    base = torch.tensor(b.reshape(-1), dtype=torch.float64, device=dev())
    s = nat.stream()
    scratch_r = torch.zeros_like(p)
    rr_tmp = torch.zeros(count * nvec, dtype=torch.float64, device=dev())

    def first_pass(pv, t):
        A.cg_update(pv, scratch_r, nvec, base.clone(), rr_tmp, passes=1, t=t)
        return A.energy_partials(nvec)[:count * nvec].clone()

    t0 = torch.empty_like(p)
    first_pass(p, t0)
    p_ref, t_ref = p.clone(), torch.empty_like(p)
    nat.call("wpc_cg_direction", n, nvec, nat.ptr(z), nat.ptr(p_ref), nat.code(dtype),
             nat.ptr(base), s)
    e_ref = first_pass(p_ref, t_ref)
    p_rec, st_rec, t_rec = p.clone(), base.clone(), t0.clone()
    A.direction_energy(z, p_rec, t_rec, nvec, 0, None, None, st_rec)
    e_rec = A.energy_partials(nvec)[:count * nvec].clone()
    torch.cuda.synchronize()

    np.testing.assert_array_equal(host(st_rec), host(base))
    np.testing.assert_array_equal(host(p_rec), host(p_ref))
    tr, tref = host(t_rec).reshape(n, nvec), host(t_ref).reshape(n, nvec)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    for v in range(nvec):
        if v == inactive:
            np.testing.assert_array_equal(tr[:, v], host(t0).reshape(n, nvec)[:, v])
            continue
        err = np.linalg.norm(tr[:, v].astype(np.float64) - tref[:, v]) / np.linalg.norm(tref[:, v])
        assert err <= tol, (v, err)
    er = host(e_rec).reshape(count, nvec).sum(axis=0)
    ef = host(e_ref).reshape(count, nvec).sum(axis=0)
    np.testing.assert_allclose(er, ef, rtol=tol)


@pytest.mark.parametrize("weight", [0.0, 1.0])
def test_recurrence_with_scalar_folds_matches_the_direction_step(weight, monkeypatch):
    import warp_preconditioners as wp
    from warp_preconditioners import linear

    monkeypatch.setattr(linear, "FOLD_IN_PLACE", 1)  # every count takes the scalar folds
    n, A, q = _operator(44, 64, torch.float64, weight)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=torch.float32)
    B = _vec(np.random.default_rng(29).normal(size=(n, 3)), torch.float64)
    calls = _counted(A)
    out = {}
    for rec in (False, True):
        work = linear.KrylovWork("cg", A, P, n, 3, torch.float64, dev())
        work.recurrence = rec
        before = len(calls)
        X = torch.zeros_like(B)
        o = linear.solve_batched("cg", A, B, X, M=P, tol=1e-10, maxiter=20000, check_every=25,
                                 work=work)
        assert (len(calls) > before) == rec
        out[rec] = (o["iterations"], host(X))
    Bh = host(B)
    for rec, (its, X) in out.items():
        for v in range(3):
            res = np.linalg.norm(Bh[:, v] - q @ X[:, v]) / np.linalg.norm(Bh[:, v])
            assert res <= 1.5e-10, (rec, v, res)
    for a, b in zip(out[True][0], out[False][0]):
        assert abs(a - b) <= max(2, 0.02 * b), (out[True][0], out[False][0])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_passes_on_the_stencil_form_equal_the_plain_slices(dtype, monkeypatch):
    """The factored operator's fused passes on the stencil form of L (no column
    stream) against the same passes on the plain slices: bit for bit."""
    from warp_preconditioners import _native as nat
    from warp_preconditioners.sparse import StreamPlan

    nvec = 3
    outs = []
    for stencil in (True, False):
        monkeypatch.setattr(StreamPlan, "STENCIL", stencil)
        n, A, _ = _operator(41, 77, dtype, 1.0)
        count = A.cg_update_parts(nvec, dtype)
        A._stream()._sliced()
        assert (int(A._stream()._sell[6]) == StreamPlan.COL_STENCIL) == stencil
        rng = np.random.default_rng(78)
        p = _vec(rng.normal(size=n * nvec), dtype)
        z = _vec(rng.normal(size=n * nvec), dtype)
        r = _vec(rng.normal(size=n * nvec), dtype)
        x = _vec(rng.normal(size=n * nvec), dtype)
        t = torch.empty_like(p)
        st = _state(nvec, 1)
        rr = torch.zeros(count * nvec, dtype=torch.float64, device=dev())
        A.cg_update(p, r, nvec, st, rr, x=x, t=t)
        part_rz = _vec(rng.uniform(0.5, 1.5, size=300 * nvec) / 300, torch.float64)
        A.direction_energy(z, p, t, nvec, 300, part_rz, rr, st)
        A.cg_update(p, r, nvec, st, rr, passes=2, x=x, t=t)
        torch.cuda.synchronize()
        outs.append([host(v) for v in (p, t, r, x, st, rr)])
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)
