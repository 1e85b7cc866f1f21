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


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("order", ["natural", "by-length"])
def test_sliced_matches_oracle_bitwise(oracle, monkeypatch, dtype, order):
    # The sliced mapping sums every row serially in storage order, long rows
    # included: bit-identical to the oracle's serial product, also with the rows
    # ordered by length inside windows (ragged last window, empty rows).
    from warp_preconditioners.sparse import StreamPlan

    monkeypatch.setattr(StreamPlan, "SORT_GAIN", -1.0 if order == "by-length" else 1e9)
    monkeypatch.setattr(StreamPlan, "SORT_WINDOW", 200)
    rng = np.random.default_rng(5)
    a = ss.random(3001, 3001, density=0.004, random_state=rng, format="lil")
    a[0, :] = 0
    a[17, :] = rng.normal(size=3001)
    a[2999, :] = 0
    a = a.tocsr()
    npd = np.float32 if dtype == torch.float32 else np.float64
    A = upload(a, dtype)
    plan = StreamPlan(A, lanes=StreamPlan.SLICED)
    x = rng.normal(size=3001).astype(npd)
    y = rng.normal(size=3001).astype(npd)
    z = torch.empty(3001, dtype=dtype, device=dev())
    plan.run(torch.tensor(x, device=dev()), torch.tensor(y, device=dev()), z, 0.7, -0.3)
    assert (plan._sell[5] is not None) == (order == "by-length")
    np.testing.assert_array_equal(host(z), oracle.spmv(oracle_csr(oracle, a, npd), x, y, 0.7, -0.3))
    if order == "by-length":
        rows = host(plan._sell[5])
        np.testing.assert_array_equal(np.sort(rows), np.arange(3001))
        lengths = np.diff(a.indptr)[rows]
        for w in range(0, 3001, 200):
            assert np.all(np.diff(lengths[w:w + 200]) <= 0)  # longest first in each window


@pytest.mark.parametrize("nvec", [1, 3, 5, 8])
@pytest.mark.parametrize("vals", [torch.float64, torch.float32, torch.bfloat16])
@pytest.mark.parametrize("sort", [False, True])
def test_sliced_equals_streamed_with_epilogue(monkeypatch, nvec, vals, sort):
    from warp_preconditioners.sparse import BsrMatrix, StreamPlan

    monkeypatch.setattr(StreamPlan, "SORT_GAIN", -1.0 if sort else 1e9)
    monkeypatch.setattr(StreamPlan, "SORT_WINDOW", 128)

    rng = np.random.default_rng(nvec)
    n = 2 * 64 * 7 + 13  # ragged last slice
    a = ss.random(n, n, density=0.01, random_state=rng, format="lil")
    for r in (3, 64, 65, 700):
        a[r, :] = 0  # empty rows inside slices
    a[100, : 150] = rng.normal(size=150)  # one long row pads its slice
    a = a.tocsr()
    A64 = upload(a)
    A = BsrMatrix(n, n, (1, 1), A64.offsets, A64.columns, A64.values.to(vals))
    xd = torch.float32 if vals != torch.float64 else torch.float64
    X = torch.tensor(rng.normal(size=(n, nvec)), dtype=xd, device=dev())
    Y = torch.tensor(rng.normal(size=(n, nvec)), dtype=xd, device=dev())
    U = torch.tensor(rng.normal(size=(n, nvec)), dtype=xd, device=dev())
    S = torch.tensor(rng.uniform(0.5, 2, size=n), dtype=torch.float32, device=dev())
    D = torch.tensor(rng.uniform(0.5, 2, size=n), dtype=xd, device=dev())
    outs = []
    for mode in (0, StreamPlan.SLICED):
        plan = StreamPlan(A, lanes=mode)
        Z = torch.empty_like(X)
        dots = torch.zeros(nvec, dtype=torch.float64, device=dev())
        plan.run(X.reshape(-1), Y.reshape(-1), Z.reshape(-1), 0.9, -0.4, nvec, s=S, d=D,
                 u=U.reshape(-1), w=X.reshape(-1), out=dots)
        outs.append((host(Z), host(dots)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-10)


def test_sliced_copy_follows_refit():
    # FSAI.update rewrites G/GT in place: the sliced copies must follow.
    from warp_preconditioners import FSAI
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(12)
    a1 = random_spd(rng, 2000, 0.004)
    a2 = a1.copy()
    nz = np.nonzero(a2)
    a2[nz] *= 1.0 + 0.1 * rng.uniform(size=len(nz[0]))
    a2 = 0.5 * (a2 + a2.T) + 2.0 * np.eye(2000)
    A1, A2 = upload(a1), upload(a2)
    P = FSAI(A1, max_row_size=8, reuse_pattern=True)
    for plan in P.plans():
        plan.fixed = StreamPlan.SLICED
    x = torch.tensor(rng.normal(size=2000), device=dev())
    z = torch.empty_like(x)
    P.apply_batch(x, x, z, 1.0, 0.0, 1)  # builds the sliced copies
    P.update(A2)
    P.apply_batch(x, x, z, 1.0, 0.0, 1)
    ref = FSAI(A1, max_row_size=8, reuse_pattern=True).update(A2)
    for plan in ref.plans():
        plan.fixed = 0  # streamed: the same summation order as the sliced mapping
    zr = torch.empty_like(x)
    ref.apply_batch(x, x, zr, 1.0, 0.0, 1)
    np.testing.assert_array_equal(host(z), host(zr))


@pytest.mark.parametrize("far", [False, "wide", True])
def test_sliced_int16_columns_and_fallback(oracle, far):
    # Columns are stored as 16-bit offsets from each slice's smallest column when a
    # slice spans at most 65,534 columns (a banded matrix; "wide": every row also
    # reaches 33,000 columns back, beyond a signed 16-bit difference from the row,
    # as the second-neighbour planes of a 128^3 mesh do); one entry 69,999 columns
    # from its row keeps the int32 columns.  All give the oracle's product bit for
    # bit.
    from warp_preconditioners.sparse import StreamPlan

    n = 70000
    rng = np.random.default_rng(21)
    off = rng.integers(-300, 300, size=(n, 5))
    rows = np.repeat(np.arange(n), 5)
    cols = np.clip(rows + off.ravel(), 0, n - 1)
    if far == "wide":
        rows = np.concatenate([rows, np.arange(33000, n)])
        cols = np.concatenate([cols, np.arange(33000, n) - 33000])
    a = ss.csr_matrix((rng.normal(size=rows.size), (rows, cols)), shape=(n, n))
    a = a.tolil()
    a[17, :] = 0  # an empty row
    if far is True:
        a[3, n - 1] = 2.5
    a = a.tocsr()
    a.eliminate_zeros()
    A = upload(a)
    plan = StreamPlan(A, lanes=StreamPlan.SLICED)
    x = rng.normal(size=(n, 2))
    z = torch.empty(n * 2, dtype=torch.float64, device=dev())
    plan.run(torch.tensor(x.ravel(), device=dev()), None, z, 1.0, 0.0, 2)
    assert plan._sell[6] == (far is not True)
    ref = np.stack([oracle.spmv(oracle_csr(oracle, a), np.ascontiguousarray(x[:, v]))
                    for v in range(2)], axis=1)
    np.testing.assert_array_equal(host(z).reshape(n, 2), ref)


@pytest.mark.parametrize("col16", [True, False])
def test_sliced_every_slice_length_bitwise(oracle, monkeypatch, col16):
    # Slices of every length 0..24 (the row in each slice's first position is the
    # longest, the others random and shorter): every remainder of the entry groups
    # (full groups of 8, masked groups of 8 and 4, groups of 4, 2, 1) against the
    # oracle's serial product, bit for bit, on int16 and int32 columns; with three
    # right-hand sides and the fused epilogue against the streamed mapping.
    from warp_preconditioners.sparse import StreamPlan

    monkeypatch.setattr(StreamPlan, "COL16", col16)
    monkeypatch.setattr(StreamPlan, "SORT_GAIN", 1e9)
    rng = np.random.default_rng(24)
    n = 25 * 64 + 13
    rows, cols = [], []
    for r in range(n):
        top = (r // 64) % 25
        m = top if r % 64 == 0 else int(rng.integers(0, top + 1))
        lo, hi = max(0, r - 300), min(n, r + 300)
        rows += [r] * m
        cols += sorted(rng.choice(np.arange(lo, hi), size=m, replace=False).tolist())
    a = ss.csr_matrix((rng.normal(size=len(rows)), (rows, cols)), shape=(n, n))
    a.sort_indices()
    A = upload(a)
    plan = StreamPlan(A, lanes=StreamPlan.SLICED)
    x = rng.normal(size=n)
    y = rng.normal(size=n)
    z = torch.empty(n, dtype=torch.float64, device=dev())
    plan.run(torch.tensor(x, device=dev()), torch.tensor(y, device=dev()), z, 0.7, -0.3)
    assert bool(plan._sell[6]) == col16
    lens = np.abs(host(plan._sliced()[1]))
    assert set(range(25)) <= set(lens.tolist())
    np.testing.assert_array_equal(host(z), oracle.spmv(oracle_csr(oracle, a, np.float64), x, y, 0.7, -0.3))
    k = 3
    X = torch.tensor(rng.normal(size=(n, k)), device=dev())
    U = torch.tensor(rng.normal(size=(n, k)), device=dev())
    S = torch.tensor(rng.uniform(0.5, 2, size=n), dtype=torch.float32, device=dev())
    D = torch.tensor(rng.uniform(0.5, 2, size=n), device=dev())
    outs = []
    for mode in (0, StreamPlan.SLICED):
        p = StreamPlan(A, lanes=mode)
        Z = torch.empty_like(X)
        dots = torch.zeros(k, dtype=torch.float64, device=dev())
        p.run(X.reshape(-1), None, Z.reshape(-1), 0.9, 0.0, k, s=S, d=D, u=U.reshape(-1),
              w=X.reshape(-1), out=dots)
        outs.append((host(Z), host(dots)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-12)


@pytest.mark.parametrize("lanes", [2, 4, 8])
@pytest.mark.parametrize("nvec", [3, 8])
def test_lanes_mapping_batched_epilogue(lanes, nvec):
    # The multi-lane mapping (two entries per lane per round trip, butterfly sums
    # with DPP quad steps) on several vectors with the fused epilogue and dot,
    # against the streamed mapping's serial sums (rows reordered: tolerance).
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(lanes * 10 + nvec)
    n = 1500
    a = ss.random(n, n, density=0.01, random_state=rng, format="lil")
    a[7, :] = 0
    a[9, : 300] = rng.normal(size=300)  # one long row
    a = a.tocsr()
    A = upload(a)
    X = torch.tensor(rng.normal(size=(n, nvec)), device=dev())
    Y = torch.tensor(rng.normal(size=(n, nvec)), device=dev())
    U = torch.tensor(rng.normal(size=(n, nvec)), device=dev())
    S = torch.tensor(rng.uniform(0.5, 2, size=n), dtype=torch.float32, device=dev())
    D = torch.tensor(rng.uniform(0.5, 2, size=n), device=dev())
    outs = []
    for mode in (0, lanes):
        p = StreamPlan(A, lanes=mode)
        Z = torch.empty_like(X)
        dots = torch.zeros(nvec, dtype=torch.float64, device=dev())
        p.run(X.reshape(-1), Y.reshape(-1), Z.reshape(-1), 0.9, -0.4, nvec, s=S, d=D,
              u=U.reshape(-1), w=X.reshape(-1), out=dots)
        outs.append((host(Z), host(dots)))
    np.testing.assert_allclose(outs[1][0], outs[0][0], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(outs[1][1], outs[0][1], rtol=1e-10)


@pytest.mark.parametrize("col16", [True, False])
def test_sliced_padding_never_multiplies_inf_or_nan(oracle, monkeypatch, col16):
    # Padding entries of the sliced copy add nothing, not 0 * x: their gathers fall
    # outside the buffer descriptor's range and read 0.  An empty row gives
    # beta * y although x holds NaN in the columns next to it, and a row shorter
    # than its slice whose last column holds inf gives inf as the serial product
    # does (0 * inf would give NaN).  alpha != 0.
    from warp_preconditioners.sparse import StreamPlan

    monkeypatch.setattr(StreamPlan, "COL16", col16)
    monkeypatch.setattr(StreamPlan, "SORT_GAIN", 1e9)
    rng = np.random.default_rng(31)
    n = 256
    a = ss.lil_matrix((n, n))
    for r in range(n):
        for c in sorted(set(rng.integers(max(0, r - 20), min(n, r + 20), size=6).tolist())):
            a[r, c] = rng.normal()
    a[5, :] = 0                   # empty row
    a[70, :] = 0
    a[70, 71] = 1.5               # one entry; its slice's longest row has more
    a[130, :] = 0
    a[130, 0] = 2.0               # a row whose only column is 0
    a = a.tocsr()
    a.eliminate_zeros()
    A = upload(a)
    plan = StreamPlan(A, lanes=StreamPlan.SLICED)
    x = rng.normal(size=n)
    x[0] = np.nan
    x[5] = np.nan
    x[71] = np.inf
    y = rng.normal(size=n)
    z = torch.empty(n, dtype=torch.float64, device=dev())
    plan.run(torch.tensor(x, device=dev()), torch.tensor(y, device=dev()), z, 0.8, 0.5)
    assert bool(plan._sell[6]) == col16
    got = host(z)
    ref = oracle.spmv(oracle_csr(oracle, a), x, y, 0.8, 0.5)
    assert got[5] == 0.5 * y[5]
    assert got[70] == np.inf
    np.testing.assert_array_equal(got, ref)   # NaN where the serial product has NaN
    cols = host(plan._sliced()[2])
    pad = (cols == -1) if not col16 else (cols.view(np.uint16) == 0xFFFF)
    assert pad.sum() > 0                      # padding is stored with the marker


@pytest.mark.parametrize("lanes", [32, 64])
def test_fused_cg_many_lanes_partials(oracle, lanes):
    # Fused CG with the operator and both FSAI passes fixed to `lanes` lanes per
    # row: the partials of those launches (2x / 4x the workgroups of 16 lanes) fit
    # the plans' buffers, and the iterates follow the oracle's PCG.
    from warp_preconditioners import FSAI, SparseOperator, linear
    from warp_preconditioners.sparse import StreamPlan

    size = 24
    d = ss.diags([-np.ones(size - 1), 2 * np.ones(size), -np.ones(size - 1)], [-1, 0, 1])
    k = ss.kronsum(d, d).tocsr()
    q = (ss.eye(size * size) / (size + 1) ** 2 + k @ k * (size + 1) ** 2).tocsr()
    b = np.random.default_rng(13).normal(size=size * size)
    Q = upload(q)
    if lanes == 32:
        op = SparseOperator(Q, row_lanes=32)
        P = FSAI(Q, max_row_size=8, apply_lanes=32)
    else:
        from warp_preconditioners.linear import _MatrixOperator

        op = _MatrixOperator(Q, lanes=64)
        P = FSAI(Q, max_row_size=8)
        P._plans = (StreamPlan(P.G, 64), StreamPlan(P.GT, 64))
    for nvec in (1, 3):
        B = torch.tensor(np.tile(b[:, None], (1, nvec)), device=dev()).contiguous()
        X = torch.zeros_like(B)
        out = linear.solve_batched("cg", op, B, X, M=P, tol=1e-10, maxiter=5000, check_every=1)
        assert out["work"].use_fused
        plan = op._stream() if lanes == 32 else op._plan
        assert plan._partials.numel() >= plan.workgroups(lanes) * nvec
        for p in P.plans():
            assert p._partials is None or p._partials.numel() >= p.workgroups(lanes) * nvec
        Po = oracle.fsai(oracle_csr(oracle, q), max_row_size=8)
        xo, ito, _ = oracle.krylov("cg", b, A=oracle_csr(oracle, q), P=Po, tol=1e-10,
                                   maxiter=5000, check_every=1)
        for v in range(nvec):
            assert abs(out["iterations"][v] - ito) <= max(2, 0.02 * ito)
            np.testing.assert_allclose(host(X[:, v]), xo, rtol=1e-8, atol=1e-10 * np.abs(xo).max())


def test_graph_replay_after_scratch_grew():
    # A KrylovWork's captured chunk holds the operators' scratch addresses; after a
    # solve with more right-hand sides grew that scratch, reusing the work recaptures
    # (a stale replay would write into freed blocks that other tensors now own).
    from warp_preconditioners import FSAI, linear

    rng = np.random.default_rng(40)
    a = random_spd(rng, 3000, 0.003)
    A = upload(a)
    P = FSAI(A, max_row_size=8)
    b = torch.tensor(rng.normal(size=(3000, 1)), device=dev())
    x1 = torch.zeros_like(b)
    out1 = linear.solve_batched("cg", A, b, x1, M=P, tol=1e-10, check_every=5)
    first = host(x1).copy()
    B8 = torch.tensor(rng.normal(size=(3000, 8)), device=dev())
    linear.solve_batched("cg", A, B8, torch.zeros_like(B8), M=P, tol=1e-6, check_every=5)
    torch.cuda.synchronize()
    sentinels = [torch.full((3000 * k,), 7.0, dtype=torch.float64, device=dev()) for k in (1, 2, 4)]
    x1.zero_()
    out2 = linear.solve_batched("cg", A, b, x1, M=P, tol=1e-10, check_every=5, work=out1["work"])
    torch.cuda.synchronize()
    assert out2["iterations"] == out1["iterations"]
    np.testing.assert_array_equal(host(x1), first)
    for s in sentinels:
        assert bool((s == 7.0).all())


def _grid_laplacian(nx, ny, nz, rng):
    """7-point pattern on an nx*ny*nz grid (natural order), random values, one row
    emptied: slices share at most 7 offsets, boundary rows miss some."""
    n = nx * ny * nz
    idx = np.arange(n).reshape(nz, ny, nx)
    rows, cols = [np.arange(n)], [np.arange(n)]
    for axis, step in ((2, 1), (1, nx), (0, nx * ny)):
        lo = [slice(None)] * 3
        hi = [slice(None)] * 3
        lo[axis], hi[axis] = slice(0, -1), slice(1, None)
        a, b = idx[tuple(lo)].ravel(), idx[tuple(hi)].ravel()
        rows += [a, b]
        cols += [b, a]
    r, c = np.concatenate(rows), np.concatenate(cols)
    m = ss.csr_matrix((rng.normal(size=r.size), (r, c)), shape=(n, n)).tolil()
    m[100, :] = 0  # an empty row
    m = m.tocsr()
    m.eliminate_zeros()
    m.sort_indices()
    return m


@pytest.mark.parametrize("nvec", [1, 3, 8])
def test_stencil_form_matches_oracle_and_streamed_bitwise(oracle, nvec):
    # The stencil form (per-slice offsets from the row, per-row masks, no column
    # stream) sums each row's entries in column order like the plain slices: equal
    # to the oracle's serial product and, with the fused epilogue and dot, to the
    # streamed mapping, bit for bit.  Missing neighbours of boundary rows read
    # nothing: NaN and inf in x next to them do not leak in.
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(70 + nvec)
    a = _grid_laplacian(21, 13, 11, rng)
    n = a.shape[0]
    A = upload(a)
    plan = StreamPlan(A, lanes=StreamPlan.SLICED)
    plan._sliced()
    assert int(plan._sell[6]) == StreamPlan.COL_STENCIL
    deltas = host(plan._sell[2]).reshape(-1, 32)
    lens = host(plan._sell[1])
    assert lens.max() == 7 and (deltas[0, :lens[0]] == np.sort(deltas[0, :lens[0]])).all()
    x = rng.normal(size=(n, nvec))
    # x[row + d] for a missing neighbour of a boundary row: row 20 (x = 20, the +x
    # boundary) misses column 21 (the next grid line's first node)
    x[21, :] = np.nan
    x[22, 0] = np.inf
    y = rng.normal(size=(n, nvec))
    z = torch.empty(n * nvec, dtype=torch.float64, device=dev())
    plan.run(torch.tensor(x.ravel(), device=dev()), torch.tensor(y.ravel(), device=dev()), z,
             0.7, -0.3, nvec)
    ref = np.stack([oracle.spmv(oracle_csr(oracle, a, np.float64),
                                np.ascontiguousarray(x[:, v]), np.ascontiguousarray(y[:, v]),
                                0.7, -0.3) for v in range(nvec)], axis=1)
    got = host(z).reshape(n, nvec)
    np.testing.assert_array_equal(got, ref)
    assert np.isfinite(got[20]).all() and got[100, 0] == -0.3 * y[100, 0]
    X = torch.tensor(rng.normal(size=(n, nvec)), device=dev())
    U = torch.tensor(rng.normal(size=(n, nvec)), device=dev())
    S = torch.tensor(rng.uniform(0.5, 2, size=n), dtype=torch.float32, device=dev())
    D = torch.tensor(rng.uniform(0.5, 2, size=n), device=dev())
    outs = []
    for mode in (0, StreamPlan.SLICED):
        p = StreamPlan(A, lanes=mode)
        Z = torch.empty_like(X)
        dots = torch.zeros(nvec, dtype=torch.float64, device=dev())
        p.run(X.reshape(-1), None, Z.reshape(-1), 0.9, 0.0, nvec, s=S, d=D, u=U.reshape(-1),
              w=X.reshape(-1), out=dots)
        outs.append((host(Z), host(dots)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-12)


def test_stencil_form_only_where_slices_share_few_offsets(monkeypatch):
    # Slices whose rows share more than 32 offsets, or whose unions of offsets
    # exceed their longest rows by more than 1/64 in all, keep the plain slices;
    # the switch turns the form off.
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(5)
    n = 640
    a = ss.random(n, n, density=0.03, random_state=rng, format="csr") + ss.eye(n)
    a.sort_indices()
    plan = StreamPlan(upload(a), lanes=StreamPlan.SLICED)
    plan._sliced()
    assert int(plan._sell[6]) != StreamPlan.COL_STENCIL
    g = _grid_laplacian(16, 8, 5, rng)
    monkeypatch.setattr(StreamPlan, "STENCIL", False)
    plan = StreamPlan(upload(g), lanes=StreamPlan.SLICED)
    plan._sliced()
    assert int(plan._sell[6]) == StreamPlan.COL_U16


@pytest.mark.parametrize("side", [33, 40])
def test_stencil_form_on_served_kuhn_grids(side, monkeypatch):
    # The served meshes' K (device assembly, Kuhn grids of any side): slices that
    # span grid lines whose boundary rows miss different neighbours store a few
    # more entries in the stencil form (0.1-0.3%), which is admitted; its products
    # equal the plain slices' bit for bit.
    from warp_preconditioners import assembly
    from warp_preconditioners.sparse import StreamPlan

    v, t = assembly.jittered_grid(side, 3, dev())
    K, _ = assembly.p1_laplacian(v, t)
    n = K.nrow
    rng = np.random.default_rng(side)
    X = torch.tensor(rng.normal(size=n * 3), device=dev())
    outs = []
    for stencil in (True, False):
        monkeypatch.setattr(StreamPlan, "STENCIL", stencil)
        plan = StreamPlan(K, lanes=StreamPlan.SLICED)
        Z = torch.empty_like(X)
        dots = torch.zeros(3, dtype=torch.float64, device=dev())
        plan.run(X, None, Z, 1.0, 0.0, 3, w=X, out=dots)
        assert (int(plan._sell[6]) == StreamPlan.COL_STENCIL) == stencil
        outs.append((host(Z), host(dots)))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
