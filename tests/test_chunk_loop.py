"""Host logic of the solvers' chunk loop (launch the next chunk before reading the
previous chunk's state), driven by a scripted stand-in for the device work."""

import torch

from warp_preconditioners.linear import _chunk_loop


class _Event:
    def __init__(self, log, index):
        self.log, self.index = log, index

    def synchronize(self):
        self.log.append(("wait", self.index))


class _Work:
    """State after each step: residual rr falls by one per active step until it
    reaches the target (then the slot rests: later steps change nothing)."""

    def __init__(self, rr0, target):
        self.rr, self.target, self.steps_taken = rr0, target, 0
        self.log = []

    def run(self, xf, steps, use_graph):
        self.log.append(("run", steps, use_graph))
        for _ in range(steps):
            if self.rr > self.target:
                self.rr -= 1
                self.steps_taken += 1

    def snapshot_state(self, index):
        self.log.append(("snap", index))
        return _Event(self.log, index), torch.tensor([0.0, self.rr, self.target, 0.0,
                                                      float(self.steps_taken)])


def _check(host, _):
    return float(host[1]) <= float(host[2])


def test_pipelined_stops_at_the_converging_chunk():
    w = _Work(rr0=60, target=0)  # converges during the third chunk of 25
    at, chunks = _chunk_loop(w, None, 25, 1000, True, 2, _check)
    assert at == 75 and chunks == 4  # one chunk past the deciding one, taking no steps
    assert w.steps_taken == 60
    runs = [e for e in w.log if e[0] == "run"]
    assert runs[0] == ("run", 25, False) and all(r[2] for r in runs[1:])
    # chunk k+1 is launched before chunk k's state is waited for
    order = [e for e in w.log if e[0] in ("snap", "wait")]
    assert order[:4] == [("snap", 0), ("snap", 1), ("wait", 0), ("snap", 2)]


def test_unpipelined_matches_checking_every_chunk():
    w = _Work(rr0=60, target=0)
    at, chunks = _chunk_loop(w, None, 25, 1000, True, 1, _check)
    assert at == 75 and chunks == 3
    order = [e for e in w.log if e[0] in ("snap", "wait")]
    assert order == [("snap", 0), ("wait", 0), ("snap", 1), ("wait", 1), ("snap", 2), ("wait", 2)]


def test_maxiter_bounds_the_launched_steps():
    w = _Work(rr0=10 ** 6, target=0)
    seen = []
    at, chunks = _chunk_loop(w, None, 25, 60, True, 2, lambda h, s: seen.append(s) or False)
    assert at == 60 and chunks == 3 and w.steps_taken == 60
    assert seen == [25, 50, 60]  # every chunk checked once, the last one partial
    assert [e for e in w.log if e[0] == "run"][-1] == ("run", 10, False)
