"""Host-side parts of the serving loop: the request stream and the stopping
certificate's arithmetic."""

import numpy as np

from warp_preconditioners.refinement import field_bounds
from warp_preconditioners.serving import KINDS, ServingEngine, request_stream


def test_request_stream_matches_the_workload_spec():
    reqs = request_stream()  # 512 requests, 200/s, seed 7, sides 16..96, repeat 0.5
    assert len(reqs) == 512
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 200) < 0.15 / 200  # Poisson, mean 5 ms
    sides = np.array([r.mesh.side for r in reqs])
    assert sides.min() >= 16 and sides.max() <= 96
    assert {r.kind for r in reqs} == set(KINDS)
    assert all(1 <= r.nrhs <= 8 for r in reqs)
    seen, repeats = set(), 0
    for r in reqs:
        key = (r.mesh, r.kind)
        repeats += key in seen
        seen.add(key)
    assert 0.4 < repeats / len(reqs) < 0.6
    again = request_stream()
    assert [(r.arrival, r.mesh, r.kind, r.nrhs, r.rhs_seed) for r in again] == \
        [(r.arrival, r.mesh, r.kind, r.nrhs, r.rhs_seed) for r in reqs]


def test_field_bounds_arithmetic():
    norms = np.array([4.0, 0.0, 1.0, 16.0, 25.0, 0.0])
    b = field_bounds(norms, 3, 2.0)
    assert b[0] == 2.0 / (2.0 * 4.0)
    assert b[1] == 0.0
    assert b[2] == np.inf  # nonzero residual, zero iterate


def test_memory_estimates_scale_with_size():
    small = request_stream(count=1, sides=(16, 16))[0]
    big = request_stream(count=1, sides=(96, 96))[0]
    s_sys, _ = ServingEngine.estimate(small, "f64")
    b_sys, _ = ServingEngine.estimate(big, "f64")
    assert b_sys / s_sys == (96 / 16) ** 3


def test_width_policy_by_kind_and_side():
    import pytest

    from warp_preconditioners.serving import fsai_width

    assert fsai_width(8, "biharmonic", 96) == 8
    policy = {"laplacian": 8, "biharmonic": [(0, 8), (64, 16)]}
    assert fsai_width(policy, "laplacian", 96) == 8
    assert fsai_width(policy, "biharmonic", 63) == 8
    assert fsai_width(policy, "biharmonic", 64) == 16
    assert fsai_width(policy, "biharmonic", 96) == 16
    with pytest.raises(ValueError, match="no FSAI width"):
        fsai_width({"biharmonic": [(32, 16)]}, "biharmonic", 16)


def test_setup_worker_takes_the_smallest_pending_job_first():
    import threading

    from warp_preconditioners.serving import PrioritySetup

    gate = threading.Event()
    order = []
    w = PrioritySetup()
    first = w.submit(0.0, gate.wait)  # holds the worker while the rest queue up
    futs = [w.submit(p, order.append, name) for p, name in
            ((5.0, "large"), (1.0, "small"), (3.0, "mid"), (1.0, "small-later"))]
    gate.set()
    for f in [first, *futs]:
        f.result(timeout=10)
    assert order == ["small", "small-later", "mid", "large"]
    boom = w.submit(0.0, lambda: 1 / 0)
    import pytest

    with pytest.raises(ZeroDivisionError):
        boom.result(timeout=10)
    w.shutdown()
