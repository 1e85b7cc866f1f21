"""Grouped CG steps: several systems of uneven size advance with one launch per
pass, bit-identical to each system's own fused CG (KrylovWork.step)."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host, laplacian_2d, upload

pytestmark = pytest.mark.gpu


def _vec(a, dtype):
    return torch.tensor(np.asarray(a), dtype=dtype, device=dev())


def _factored(side, seed, factor=torch.float32):
    import warp_preconditioners as wp
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(seed)
    k = laplacian_2d(side)
    n = k.shape[0]
    mass = rng.uniform(0.5, 1.5, size=n) / side ** 2
    q = (k @ ss.diags(1 / mass) @ k + ss.diags(mass)).tocsr()
    A = wp.SquaredLaplacianOperator(upload(k), _vec(mass, torch.float64),
                                    torch.arange(n, dtype=torch.int32, device=dev()),
                                    row_lanes=1, mass_weight=1.0)
    P = wp.FSAI(upload(q), max_row_size=8, factor_dtype=factor)
    for plan in (A._stream(), *P.plans()):
        plan.fixed = StreamPlan.SLICED
    return n, A, P


def _plain(side, seed, dtype=torch.float32, factor=torch.bfloat16):
    import warp_preconditioners as wp
    from warp_preconditioners.sparse import StreamPlan

    rng = np.random.default_rng(seed)
    k = laplacian_2d(side)
    n = k.shape[0]
    a = (k + ss.diags(rng.uniform(0.5, 1.5, size=n) / side ** 2)).tocsr()
    A = wp.SparseOperator(upload(a, dtype), row_lanes=1)
    P = wp.FSAI(upload(a), max_row_size=8, factor_dtype=factor)
    for plan in (A._stream(), *P.plans()):
        plan.fixed = StreamPlan.SLICED
    return n, A, P


@pytest.mark.parametrize("case", ["factored-f64", "plain-f32-bf16"])
@pytest.mark.parametrize("nvecs", [(1, 1, 1, 1), (3, 1, 8, 2)])
def test_grouped_steps_equal_own_steps(case, nvecs):
    # uneven sizes and (second case) uneven right-hand-side counts in one batch
    from warp_preconditioners import grouped
    from warp_preconditioners.linear import KrylovWork

    dtype = torch.float64 if case == "factored-f64" else torch.float32
    systems = [(_factored if case == "factored-f64" else _plain)(side, seed)
               for side, seed in ((9, 1), (40, 2), (23, 3), (57, 4))]
    rng = np.random.default_rng(sum(nvecs))
    own, mine, xs_own, xs_mine = [], [], [], []
    for (n, A, P), nvec in zip(systems, nvecs):
        B = _vec(rng.normal(size=(n, nvec)), dtype)
        for works, xs in ((own, xs_own), (mine, xs_mine)):
            w = KrylovWork("cg", A, P, n, nvec, dtype, dev())
            w.fuse_update = False  # the grouped kernels keep the separate update step
            X = torch.zeros_like(B)
            w.start(B.reshape(-1), X.reshape(-1), 1e-9 if dtype == torch.float64 else 1e-5, 0.0)
            works.append(w)
            xs.append(X)
    keys = {grouped.group_key(w) for w in mine}
    assert len(keys) == 1 and None not in keys
    group = grouped.GroupedCG(keys.pop(), steps=7)
    side = torch.cuda.Stream()  # a capturable stream for the batch
    side.wait_stream(torch.cuda.current_stream())
    for chunk in range(6):
        # systems leave and rejoin the batch between chunks
        members = [0, 1, 2, 3] if chunk % 3 == 0 else ([1, 3] if chunk % 3 == 1 else [0, 2, 3])
        with torch.cuda.stream(side):
            group.run([mine[i] for i in members], [xs_mine[i].reshape(-1) for i in members])
        for i in members:
            own[i].run(xs_own[i].reshape(-1), 7, use_graph=False)
        torch.cuda.synchronize()
        for i in range(4):
            np.testing.assert_array_equal(host(xs_mine[i]), host(xs_own[i]))
            np.testing.assert_array_equal(host(mine[i].state), host(own[i].state))
    assert group.graph is not None  # later chunks replayed the captured graph
    group.release()
