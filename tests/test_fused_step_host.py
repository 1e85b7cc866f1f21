"""Host logic of the fused CG step sequence (linear.KrylovWork.step), driven by
stand-in operators with the launches recorded instead of run: where x is advanced
(the update on a solve's first step, the direction step once the preconditioner's
partial count is known to fold in place) and the fallbacks for large counts."""

import pytest
import torch

import warp_preconditioners._native as nat
from warp_preconditioners.linear import FOLD_IN_PLACE, KrylovWork


class _Op:
    fused = True

    def __init__(self, parts):
        self.parts = parts

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        return torch.zeros(self.parts * nvec, dtype=torch.float64), self.parts


@pytest.fixture
def launches(monkeypatch):
    log = []
    monkeypatch.setattr(nat, "call", lambda name, *args: log.append((name, args)))
    monkeypatch.setattr(nat, "stream", lambda: 0)
    return log


def _work(a_parts, m_parts, n=262144, nvec=3):
    nat.lib()  # the library loads on a host without a GPU (host-only helpers)
    return KrylovWork("cg", _Op(a_parts), _Op(m_parts), n, nvec, torch.float64, "cpu")


def test_x_moves_to_direction_after_first_step(launches):
    w = _work(1024, 1024)
    x = torch.zeros(w.n * w.nvec, dtype=torch.float64)
    for _ in range(3):
        w.step(x)
    names = [name for name, _ in launches]
    assert names == ["wpc_cg_update_fold", "wpc_cg_direction_fold"] * 3
    (_, u0), (_, d0), (_, u1), (_, d1) = launches[:4]
    assert u0[2] == x.data_ptr() and d0[4] is None   # first step: x in the update
    assert u1[2] is None and d1[4] == x.data_ptr()   # then: x in the direction step
    (_, u2), (_, d2) = launches[4:]
    assert u2[2] is None and d2[4] == x.data_ptr()


def test_large_preconditioner_counts_keep_x_in_update(launches):
    nvec = 3
    w = _work(1024, FOLD_IN_PLACE // nvec + 1, nvec=nvec)
    x = torch.zeros(w.n * nvec, dtype=torch.float64)
    for _ in range(2):
        w.step(x)
    names = [name for name, _ in launches]
    assert names == ["wpc_cg_update_fold", "wpc_cg_beta", "wpc_cg_direction"] * 2
    assert all(args[2] == x.data_ptr() for name, args in launches
               if name == "wpc_cg_update_fold")


def test_large_operator_counts_use_scalar_kernels(launches):
    nvec = 3
    w = _work(FOLD_IN_PLACE // nvec + 1, 1024, nvec=nvec)
    x = torch.zeros(w.n * nvec, dtype=torch.float64)
    for _ in range(2):
        w.step(x)
    names = [name for name, _ in launches]
    assert names == ["wpc_cg_alpha", "wpc_cg_update", "wpc_cg_beta", "wpc_cg_direction"] * 2


def test_partials_sized_for_the_launched_mapping():
    # A plan fixed to 32 (or 64) lanes per row launches ceil(32 n / 256) workgroups,
    # twice those of 16 lanes: the fused-dot partials must hold one per workgroup
    # and vector of that launch.
    import numpy as np
    import scipy.sparse as ss

    from warp_preconditioners.sparse import StreamPlan, from_scipy

    nat.lib()
    a = ss.random(5000, 5000, density=0.002, random_state=np.random.default_rng(3), format="csr")
    A = from_scipy(a + ss.eye(5000), "cpu")
    for fixed in (None, 16, 32, 64):
        plan = StreamPlan(A, lanes=fixed)
        for nvec in (1, 3, 8):
            need = plan.partials_size(nvec, fixed)
            launched = plan.workgroups(plan.SLICED if fixed is None else fixed)
            assert need >= launched * nvec
            assert need >= max(plan.workgroups(m) for m in plan.CANDIDATES) * nvec
    plan = StreamPlan(A, lanes=32)
    assert plan.partials_size(1, 32) == plan.workgroups(32) == -(-5000 * 32 // 256)


def test_graph_replay_keyed_on_scratch_generation(monkeypatch):
    # KrylovWork replays a captured chunk only while x and every operator scratch
    # buffer are the ones it captured; a reallocation (generation bump) recaptures.
    import warp_preconditioners.linear as lin

    captured = []

    class _Graph:
        def replay(self):
            captured.append("replay")

    class _Ctx:
        def __init__(self, g):
            captured.append("capture")

        def __enter__(self):
            return self

        def __exit__(self, *exc):
            return False

    monkeypatch.setattr(lin.torch.cuda, "CUDAGraph", _Graph)
    monkeypatch.setattr(lin.torch.cuda, "graph", _Ctx)
    w = _work(64, 64)
    monkeypatch.setattr(w, "step", lambda x: None)
    x = torch.zeros(8, dtype=torch.float64)
    w.run(x, 25, True)
    w.run(x, 25, True)
    assert captured == ["capture", "replay", "replay"]
    nat.bump_generation()  # e.g. the preconditioner's temporary grew for more RHS
    w.run(x, 25, True)
    assert captured == ["capture", "replay", "replay", "capture", "replay"]
    w.run(x, 25, True)
    assert captured[-1] == "replay" and captured.count("capture") == 2
    w.run(torch.zeros(8, dtype=torch.float64), 25, True)  # another x: recapture
    assert captured.count("capture") == 3


class _LogOp(_Op):
    """The factored operator's fused pair, recorded (in-place partial counts)."""

    def __init__(self, parts, log, tag):
        super().__init__(parts)
        self.log, self.tag = log, tag

    def apply_batch(self, x, y, z, alpha, beta, nvec, dot_w=None, dot_out=None):
        self.log.append((self.tag, ()))
        return super().apply_batch(x, y, z, alpha, beta, nvec, dot_w, dot_out)

    def cg_update_parts(self, nvec, dtype):
        return self.parts

    def cg_update(self, p, r, nvec, state, part_rr, passes=3, x=None, t=None):
        self.log.append(("cg_update", (passes, x is not None, t)))
        return self.parts

    def direction_energy(self, z, p, t, nvec, nparts_rz, part_rz, part_rr, state):
        self.log.append(("direction_energy", (t, nparts_rz)))

    def energy_partials(self, nvec):
        return torch.zeros(self.parts * 8, dtype=torch.float64)


@pytest.mark.parametrize("recurrence", [True, False])
def test_recurrence_steps_defer_the_direction_step_inside_a_run(launches, recurrence):
    nvec = 3
    nat.lib()
    A, M = _LogOp(512, launches, "A"), _LogOp(1024, launches, "M")
    w = KrylovWork("cg", A, M, 262144, nvec, torch.float64, "cpu")
    w.recurrence = recurrence
    x = torch.zeros(w.n * nvec, dtype=torch.float64)
    w.steps(x, 1)  # the first step: the general sequence (the partial count is learnt)
    assert [n for n, _ in launches] == ["A", "wpc_cg_update_fold", "M", "wpc_cg_direction_fold"]
    del launches[:]
    w.steps(x, 4)
    names = [n for n, _ in launches]
    if recurrence:
        assert names == (["cg_update", "M"] + ["direction_energy", "cg_update", "M"] * 3
                         + ["wpc_cg_direction_fold"])
        ups = [a for n, a in launches if n == "cg_update"]
        assert [u[0] for u in ups] == [3, 2, 2, 2]          # both passes, then pass 2 alone
        assert all(u[1] and u[2] is w.ap for u in ups)       # x advanced; t carried in ap
        assert all(a == (w.ap, 1024) for n, a in launches if n == "direction_energy")
        fold = launches[-1][1]
        assert fold[4] is None and fold[6] == 1024 and fold[8] == 512
    else:
        assert names == ["cg_update", "M", "wpc_cg_direction_fold"] * 4
    # a run of one step is one whole step
    del launches[:]
    w.steps(x, 1)
    assert [n for n, _ in launches] == ["cg_update", "M", "wpc_cg_direction_fold"]


def test_recurrence_with_scalar_folds_past_the_in_place_fold(launches):
    # Partial counts too large to fold in every workgroup: the recurrence pass takes
    # BETA from the state (wpc_cg_beta each step), and only the run's last step
    # launches the direction step.
    nvec = 3
    nat.lib()
    big = FOLD_IN_PLACE // nvec + 1
    A = _LogOp(big, launches, "A")
    w = KrylovWork("cg", A, _LogOp(64, launches, "M"), 262144, nvec, torch.float64, "cpu")
    x = torch.zeros(w.n * nvec, dtype=torch.float64)
    w.steps(x, 1)
    assert [n for n, _ in launches] == ["A", "wpc_cg_alpha", "wpc_cg_update", "M",
                                        "wpc_cg_beta", "wpc_cg_direction"]
    del launches[:]
    w.steps(x, 3)
    names = [n for n, _ in launches]
    step = ["wpc_cg_alpha", "cg_update", "M", "wpc_cg_beta"]
    assert names == (["cg_update"] + step + ["direction_energy"] + step
                     + ["direction_energy"] + step + ["wpc_cg_direction"])
    ups = [a for n, a in launches if n == "cg_update"]
    assert [u[0] for u in ups] == [1, 6, 6, 6]
    assert [u[1] for u in ups] == [False, True, True, True]   # x in the second pass
    assert all(u[2] is w.ap for u in ups)
    assert all(a == (w.ap, 0) for n, a in launches if n == "direction_energy")
