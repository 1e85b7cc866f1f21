"""Certified residuals, refinement steps, native graph capture, factor refit
reuse and the serving loop (continuous batching) on the device."""

import math
from fractions import Fraction

import numpy as np
import pytest
import scipy.sparse as ss
import scipy.sparse.linalg as sla
import torch

from gpu_helpers import dev, host

pytestmark = pytest.mark.gpu


def _mesh(side, seed):
    from warp_preconditioners import assembly

    v, t = assembly.jittered_grid(side, seed, dev())
    K, m = assembly.p1_laplacian(v, t)
    return v, t, K, m


def _csr(K):
    return ss.csr_matrix((host(K.values), host(K.columns), host(K.offsets)), shape=(K.nrow, K.ncol))


def _exact_factored_residual(Kh, mass, inv_m, alpha, x, b):
    """r = b - alpha*m*x - K diag(inv_m) K x in exact rational arithmetic, rounded."""
    n = Kh.shape[0]
    F = [Fraction(float(v)) for v in x]
    Kx = [sum((Fraction(float(Kh.data[e])) * F[Kh.indices[e]]
               for e in range(Kh.indptr[i], Kh.indptr[i + 1])), Fraction(0)) for i in range(n)]
    t = [Kx[i] * Fraction(float(inv_m[i])) for i in range(n)]
    a = Fraction(float(alpha))
    r = []
    for i in range(n):
        acc = Fraction(float(b[i])) - a * Fraction(float(mass[i])) * F[i]
        for e in range(Kh.indptr[i], Kh.indptr[i + 1]):
            acc -= Fraction(float(Kh.data[e])) * t[Kh.indices[e]]
        r.append(float(acc))
    return np.array(r)


def _exact_assembled_residual(Ah, x, b):
    n = Ah.shape[0]
    out = []
    for i in range(n):
        acc = Fraction(float(b[i]))
        for e in range(Ah.indptr[i], Ah.indptr[i + 1]):
            acc -= Fraction(float(Ah.data[e])) * Fraction(float(x[Ah.indices[e]]))
        out.append(float(acc))
    return np.array(out)


@pytest.mark.parametrize("smooth", [False, True])
def test_dd_residual_factored_matches_exact(smooth):
    from warp_preconditioners import SquaredLaplacianOperator
    from warp_preconditioners.refinement import TrueResidual

    v, _, K, m = _mesh(6, 3)
    n = K.nrow
    op = SquaredLaplacianOperator(K, m, torch.arange(n, dtype=torch.int32, device=dev()),
                                  mass_weight=0.7)
    rng = np.random.default_rng(1)
    vh = host(v)
    if smooth:  # smooth fields: the plain float64 product cancels heavily
        X = np.stack([vh[:, 0] + 0.3 * vh[:, 1] ** 2, np.sin(vh[:, 2])], axis=1)
    else:
        X = rng.standard_normal((n, 2))
    Bh = rng.standard_normal((n, 2)) * 1e-3
    x = torch.tensor(X, device=dev()).reshape(-1)
    b = torch.tensor(Bh, device=dev()).reshape(-1)
    res = TrueResidual(op, m, 2)
    norms = host(res.run(x, b))
    r = host(res.r).reshape(n, 2)
    Kh, mh, im = _csr(op.L), host(m), host(op.inverse_mass)
    for c in range(2):
        exact = _exact_factored_residual(Kh, mh, im, 0.7, X[:, c], Bh[:, c])
        err = np.abs(r[:, c] - exact)
        assert np.all(err <= 4 * np.spacing(np.abs(exact)) + 1e-300), err.max()
        np.testing.assert_allclose(norms[c], np.sum(exact ** 2 / mh), rtol=1e-12)
        np.testing.assert_allclose(norms[2 + c], np.sum(mh * X[:, c] ** 2), rtol=1e-12)
    if smooth:
        # the plain float64 operator is far less accurate on the same input
        z = torch.empty_like(x)
        op.apply_batch(x, b, z, -1.0, 1.0, 2)
        plain = host(z).reshape(n, 2)[:, 0]
        exact = _exact_factored_residual(Kh, mh, im, 0.7, X[:, 0], Bh[:, 0])
        assert np.max(np.abs(plain - exact)) > 100 * np.max(np.abs(r[:, 0] - exact))


def test_dd_residual_assembled_matches_exact():
    from warp_preconditioners import SparseOperator, assembly
    from warp_preconditioners.refinement import TrueResidual, field_bounds

    _, _, K, m = _mesh(7, 5)
    n = K.nrow
    A = assembly.add_diagonal(K, m, 1.0)
    op = SparseOperator(A, row_lanes=1)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((n, 3))
    Bh = rng.standard_normal((n, 3))
    res = TrueResidual(op, m, 3)
    norms = host(res.run(torch.tensor(X, device=dev()).reshape(-1),
                         torch.tensor(Bh, device=dev()).reshape(-1)))
    r = host(res.r).reshape(n, 3)
    Ah = _csr(A)
    for c in range(3):
        exact = _exact_assembled_residual(Ah, X[:, c], Bh[:, c])
        assert np.all(np.abs(r[:, c] - exact) <= 4 * np.spacing(np.abs(exact)) + 1e-300)
    mh = host(m)
    bounds = field_bounds(norms, 3, 1.0)
    want = np.sqrt(np.sum(r ** 2 / mh[:, None], 0)) / np.sqrt(np.sum(mh[:, None] * X ** 2, 0))
    np.testing.assert_allclose(bounds, want, rtol=1e-12)


def test_field_bound_is_an_upper_bound():
    # K + M on a jittered mesh: the certified bound dominates the true field error.
    from warp_preconditioners import SparseOperator, assembly
    from warp_preconditioners.refinement import TrueResidual, field_bounds

    _, _, K, m = _mesh(8, 11)
    n = K.nrow
    A = assembly.add_diagonal(K, m, 1.0)
    Ah = _csr(A).tocsc()
    rng = np.random.default_rng(3)
    b = rng.standard_normal((n, 2))
    xs = np.column_stack([sla.spsolve(Ah, b[:, c]) for c in range(2)])
    x = xs + 1e-6 * rng.standard_normal(xs.shape) * np.abs(xs).max()
    res = TrueResidual(SparseOperator(A, row_lanes=1), m, 2)
    norms = host(res.run(torch.tensor(x, device=dev()).reshape(-1),
                         torch.tensor(b, device=dev()).reshape(-1)))
    bound = field_bounds(norms, 2, 1.0)
    mh = host(m)[:, None]
    actual = np.sqrt(np.sum(mh * (x - xs) ** 2, 0) / np.sum(mh * xs ** 2, 0))
    assert np.all(actual <= bound * (1 + 1e-6)), (actual, bound)


def test_witness_bound_is_an_upper_bound():
    # Stiff factored biharmonic with a manufactured solution: b = A x* is formed in
    # compensated arithmetic, so x* is exact to ~1e-14.  For x' = fl(x + c) with an
    # accurate correction c, the witness certificate dominates the true field error
    # and stays near 2 eps, while the lambda_min bound of x' is held up by the
    # rounding of x' itself.
    from warp_preconditioners import SquaredLaplacianOperator, workloads
    from warp_preconditioners.refinement import TrueResidual, field_bounds, witness_bounds

    v, _, K, m = _mesh(14, 4)
    n = K.nrow
    op = SquaredLaplacianOperator(K, m, torch.arange(n, dtype=torch.int32, device=dev()),
                                  mass_weight=1.0)
    vh = host(v)
    xs = np.cos(np.pi * vh[:, 0]) + vh[:, 1] * vh[:, 2]
    res = TrueResidual(op, m, 1)
    zero = torch.zeros(n, dtype=torch.float64, device=dev())
    res.run(torch.tensor(-xs, device=dev()), zero)
    b = res.r.clone()  # = A x*
    rng = np.random.default_rng(8)
    x = xs + 1e-6 * np.sin(3 * vh[:, 1]) + 1e-9 * rng.standard_normal(n)
    prev = host(res.run(torch.tensor(x, device=dev()), b)).copy()
    r = host(res.r).copy()
    Qh = workloads.squared_operator(_csr(op.L), host(m), 1.0).tocsc()
    c = sla.splu(Qh).solve(r)
    w = torch.empty(2, dtype=torch.float64, device=dev())
    res.run(torch.tensor(c, device=dev()), res.r, norms=w)
    x1 = x + c
    nxt = host(res.run(torch.tensor(x1, device=dev()), b)).copy()
    wb = witness_bounds(nxt, host(w), prev, 1, 1.0)[0]
    pb = field_bounds(nxt, 1, 1.0)[0]
    mh = host(m)
    actual = np.sqrt(np.sum(mh * (x1 - xs) ** 2) / np.sum(mh * xs ** 2))
    assert actual <= wb + 1e-13, (actual, wb)
    assert wb < 1e-11 and pb > 10 * wb, (wb, pb)


def test_refine_prepare_update():
    from warp_preconditioners import _native as nat

    n, k = 1000, 3
    rng = np.random.default_rng(4)
    r = rng.standard_normal(n * k)
    norms = np.array([4.0, 0.0, 9.0, 1.0, 1.0, 1.0])
    rd = torch.tensor(r, device=dev())
    nd = torch.tensor(norms, device=dev())
    out = torch.empty(n * k, dtype=torch.float32, device=dev())
    mask = 0b101
    nat.call("wpc_refine_prepare", n, k, nat.ptr(rd), nat.ptr(nd), mask, nat.ptr(out),
             nat.code(torch.float32), nat.stream())
    scale = np.array([2.0, 1.0, 3.0])
    want = (r.reshape(n, k) / scale).astype(np.float32)
    want[:, 1] = 0
    np.testing.assert_array_equal(host(out).reshape(n, k), want)
    x = torch.tensor(rng.standard_normal(n * k), device=dev())
    x0 = host(x).reshape(n, k).copy()
    c64 = torch.full((n * k,), 7.0, dtype=torch.float64, device=dev())
    nat.call("wpc_refine_update", n, k, nat.ptr(x), nat.ptr(out), nat.code(torch.float32),
             nat.ptr(nd), mask, nat.ptr(c64), nat.stream())
    d = scale * want.astype(np.float64)
    d[:, 1] = 0.0
    np.testing.assert_array_equal(host(c64).reshape(n, k), d)
    np.testing.assert_array_equal(host(x).reshape(n, k), x0 + d)


def test_native_graph_capture_replays():
    import ctypes

    from warp_preconditioners import _native as nat

    n = 4096
    x = torch.arange(n, dtype=torch.float64, device=dev())
    y = torch.ones(n, dtype=torch.float64, device=dev())
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        nat.call("wpc_graph_begin", nat.stream())
        nat.call("wpc_axpby", n, 2.0, nat.ptr(x), 1.0, nat.ptr(y), nat.ptr(y),
                 nat.code(torch.float64), nat.stream())
        handle = ctypes.c_void_p()
        assert nat.lib().wpc_graph_end(nat.stream(), ctypes.addressof(handle)) == 0
        assert torch.equal(y, torch.ones_like(y))  # capture does not execute
        for rep in range(1, 4):
            nat.call("wpc_graph_launch", handle.value, nat.stream())
            s.synchronize()
            assert torch.equal(y, 1.0 + 2.0 * rep * x)
        nat.call("wpc_graph_destroy", handle.value)


def test_private_capture_streams():
    # torch hands out streams from a pool of 32 per device, so 40 torch streams
    # repeat; the capture streams come from the library, differ from every pooled
    # stream and per thread, and a graph recorded on one replays on another stream.
    import ctypes
    import threading

    from warp_preconditioners import _native as nat

    pooled = {torch.cuda.Stream().cuda_stream for _ in range(40)}
    mine = nat.capture_stream()
    assert mine is nat.capture_stream()
    other = {}
    t = threading.Thread(target=lambda: other.setdefault("s", nat.capture_stream().cuda_stream))
    t.start()
    t.join()
    assert mine.cuda_stream not in pooled
    assert other["s"] not in pooled and other["s"] != mine.cuda_stream
    n = 1000
    x = torch.arange(n, dtype=torch.float64, device=dev())
    y = torch.ones(n, dtype=torch.float64, device=dev())
    run = torch.cuda.Stream()
    handle = ctypes.c_void_p()
    with torch.cuda.stream(mine):
        nat.call("wpc_graph_begin", nat.stream())
        nat.call("wpc_axpby", n, 2.0, nat.ptr(x), 1.0, nat.ptr(y), nat.ptr(y),
                 nat.code(torch.float64), nat.stream())
        assert nat.lib().wpc_graph_end(nat.stream(), ctypes.addressof(handle)) == 0
    for rep in range(1, 4):
        nat.call("wpc_graph_launch", handle.value, run.cuda_stream)
        run.synchronize()
        assert torch.equal(y, 1.0 + 2.0 * rep * x)
    nat.call("wpc_graph_destroy", handle.value)
    extra = nat.PrivateStream()
    assert extra.handle not in pooled and extra.handle != mine.cuda_stream
    extra.release()


@pytest.mark.parametrize("factor", [torch.float32, torch.bfloat16])
def test_refit_clone_equals_update(factor):
    from warp_preconditioners import FSAI, assembly
    from warp_preconditioners.fsai_reuse import refit_clone

    _, _, K1, m1 = _mesh(9, 1)
    _, _, K2, m2 = _mesh(9, 2)
    Q1 = assembly.squared_operator(K1, m1, 1.0)
    Q2 = assembly.squared_operator(K2, m2, 1.0)
    template = FSAI(Q1, max_row_size=8, factor_dtype=factor)
    clone = refit_clone(template, Q2)
    ref = FSAI(Q1, max_row_size=8, factor_dtype=factor, reuse_pattern=True).update(Q2)
    for a, b in ((clone.G, ref.G), (clone.GT, ref.GT)):
        assert torch.equal(a.offsets, b.offsets) and torch.equal(a.columns, b.columns)
        assert torch.equal(a.values, b.values)
    if factor == torch.bfloat16:
        assert torch.equal(clone.row_scale2, ref.row_scale2)
    # the template is untouched and the clone shares its index buffers
    assert torch.equal(template.G.values, FSAI(Q1, max_row_size=8, factor_dtype=factor).G.values)
    assert clone.G.offsets.data_ptr() == template.G.offsets.data_ptr()
    with pytest.raises(ValueError, match="topology|sparsity"):
        _, _, K3, m3 = _mesh(8, 1)
        refit_clone(template, assembly.squared_operator(K3, m3, 1.0))


def _ld_matvec(K, x):
    """K x in extended precision (numpy longdouble: 64-bit significand on x86-64)."""
    prod = K.data.astype(np.longdouble) * x[K.indices]
    out = np.zeros(K.shape[0], dtype=np.longdouble)
    nz = np.diff(K.indptr) > 0
    out[nz] = np.add.reduceat(prod, K.indptr[:-1][nz])
    return out


def _reference_solution(req):
    """Host direct solve of one request (the device-assembled K and M, the same
    right-hand sides), refined with residuals of the served operator itself
    (K + alpha M, or the factored alpha M + K M^-1 K) in extended precision: a
    plain float64 direct solve of the biharmonic Q is only ~kappa * eps accurate
    (a few 1e-9 at side 28), coarser than the certificates it must check."""
    from warp_preconditioners import workloads

    _, _, Kd, md = _mesh(req.mesh.side, req.mesh.seed)
    K, m = _csr(Kd), host(md)
    K.sort_indices()
    a = req.alpha
    if req.kind == "laplacian":
        A = (K + a * ss.diags(m)).tocsc()
    else:
        A = workloads.squared_operator(K, m, a).tocsc()
    ml = m.astype(np.longdouble)

    def apply_ld(x):
        x = x.astype(np.longdouble)
        if req.kind == "laplacian":
            return _ld_matvec(K, x) + a * ml * x
        return a * ml * x + _ld_matvec(K, _ld_matvec(K, x) / ml)

    gen = torch.Generator(device=dev())
    gen.manual_seed(req.rhs_seed)
    B = host(torch.randn((len(m), req.nrhs), generator=gen, dtype=torch.float64, device=dev()))
    lu = sla.splu(A)
    out = []
    for c in range(req.nrhs):
        b = B[:, c].astype(np.longdouble)
        x = lu.solve(B[:, c]).astype(np.longdouble)
        for _ in range(3):
            x = x + lu.solve((b - apply_ld(x)).astype(np.float64))
        out.append(x.astype(np.float64))
    return np.column_stack(out), m


def _field_error(x, xs, m):
    return np.sqrt(np.sum(m[:, None] * (x - xs) ** 2, 0) / np.sum(m[:, None] * xs ** 2, 0))


@pytest.mark.parametrize("group", [False, True])
def test_serving_results_certified_and_mix_invariant(group):
    # With group=True same-shape solves in flight advance in batched launches
    # (grouped.GroupedCG); results stay bit-identical to solving alone.
    from warp_preconditioners.serving import ServingConfig, ServingEngine, request_stream

    reqs = request_stream(count=14, rate=2000.0, seed=7, sides=(6, 13), repeat=0.5, max_rhs=4)
    assert {r.kind for r in reqs} == {"laplacian", "biharmonic"}
    cfg = ServingConfig(streams=3, max_active=8, chunk=10, group=group)
    eng = ServingEngine(cfg)
    mix = eng.serve(reqs, realtime=True)
    assert (eng.stats["grouped_chunks"] > 0) == group
    provenance = {r.factor for r in mix}
    assert "cached" in provenance and "built" in provenance
    for req, res in zip(reqs, mix):
        assert res.converged and max(res.field_bound) <= 1e-8, (req, res.field_bound)
        xs, m = _reference_solution(req)
        err = _field_error(res.x, xs, m)
        assert np.all(err <= 1e-8), (req, err)
        assert np.all(err <= np.array(res.field_bound) * 1.01 + 1e-10)
        # alone under the same configuration (grouping also decides whether the
        # update step is folded into the operator's pass)
        alone = ServingEngine(ServingConfig(streams=1, max_active=1, chunk=10, group=group)).serve(
            [req], realtime=False)[0]
        if res.factor in ("built", "cached"):
            np.testing.assert_array_equal(alone.x, res.x)
            assert alone.iterations == res.iterations
        else:
            assert np.all(_field_error(res.x, alone.x, m) <= 2e-8)


def test_serving_budget_evicts_and_still_serves():
    from warp_preconditioners.serving import MeshSpec, Request, ServingConfig, ServingEngine

    # three meshes of one size, a budget of one system's (conservative) estimate:
    # every switch of mesh must evict the idle previous system first
    reqs = [Request(i, 0.0, MeshSpec(11, 100 + i % 3), "laplacian", 2, 50 + i) for i in range(6)]
    one = sum(ServingEngine.estimate(reqs[0], "mixed"))
    cfg = ServingConfig(budget_bytes=1.05 * one, streams=2, max_active=4, chunk=10)
    eng = ServingEngine(cfg)
    out = eng.serve(reqs, realtime=False)
    assert all(r.converged for r in out)
    assert eng.stats["evicted"] > 0 and eng.stats["over_budget"] == 0
    assert eng.stats["peak_bytes"] <= cfg.budget_bytes
    for req, res in zip(reqs, out):
        xs, m = _reference_solution(req)
        assert np.all(_field_error(res.x, xs, m) <= 1e-8)


def test_serving_configuration_sides_certified_and_mix_invariant():
    # Meshes of the configuration-5 range (sides 16-28 here, where host direct solves
    # stay quick), both kinds, built / cached / refit factors all present: every
    # result certified to 1e-8 and within it of a host direct solve, and equal to
    # the same request solved alone (bit for bit for built and cached factors).
    from warp_preconditioners.serving import MeshSpec, Request, ServingConfig, ServingEngine

    plan = [(16, 11, "laplacian", 3), (20, 12, "biharmonic", 2), (16, 11, "laplacian", 5),
            (16, 13, "laplacian", 1), (20, 12, "biharmonic", 4), (20, 14, "biharmonic", 3),
            (28, 15, "biharmonic", 2), (24, 16, "laplacian", 8)]
    reqs = [Request(i, 0.001 * i, MeshSpec(side, seed), kind, nrhs, 900 + i)
            for i, (side, seed, kind, nrhs) in enumerate(plan)]
    eng = ServingEngine(ServingConfig(streams=3, max_active=8, chunk=25))
    mix = eng.serve(reqs, realtime=True)
    kinds = {(r.kind, res.factor) for r, res in zip(reqs, mix)}
    for kind in ("laplacian", "biharmonic"):
        assert {(kind, "built"), (kind, "cached"), (kind, "refit")} <= kinds, kinds
    for req, res in zip(reqs, mix):
        assert res.converged and max(res.field_bound) <= 1e-8, (req, res.field_bound)
        xs, m = _reference_solution(req)
        err = _field_error(res.x, xs, m)
        assert np.all(err <= 1e-8), (req, err)
        assert np.all(err <= np.array(res.field_bound) * 1.01 + 1e-10)
        alone = ServingEngine(ServingConfig(streams=1, max_active=1, chunk=25)).serve(
            [req], realtime=False)[0]
        if res.factor in ("built", "cached"):
            np.testing.assert_array_equal(alone.x, res.x)
            assert alone.iterations == res.iterations
        else:
            assert np.all(_field_error(res.x, alone.x, m) <= 2e-8)


def test_serving_deadline_stops_and_reports_unfinished():
    # A stream cut by the wall-clock guard returns None for the requests it did not
    # finish and flags the engine as timed out (the bench then reports no rate).
    from warp_preconditioners.serving import MeshSpec, Request, ServingConfig, ServingEngine

    reqs = [Request(i, 0.0, MeshSpec(10, 300 + i), "laplacian", 2, 40 + i) for i in range(3)]
    eng = ServingEngine(ServingConfig(streams=2, max_active=4, chunk=10))
    out = eng.serve(reqs, realtime=False, deadline=0.0)
    assert eng.timed_out and out == [None, None, None]
    eng = ServingEngine(ServingConfig(streams=2, max_active=4, chunk=10))
    out = eng.serve(reqs, realtime=False, deadline=600.0)
    assert not eng.timed_out and all(r is not None and r.converged for r in out)
