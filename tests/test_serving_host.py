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
