"""Continuous-batching solve service on one GPU.

A request names a mesh (grid side and jitter seed of the unit-cube Kuhn mesh), a
system kind and 1..8 right-hand sides:

* ``"laplacian"`` -- (K + alpha*M) x = b, assembled, FSAI width 8 on it;
* ``"biharmonic"`` -- (alpha*M + K M^-1 K) x = b, applied factored (the reference's
  SquaredLaplacianOperator path), FSAI width 8 on the assembled Q.

The engine keeps *systems* (mesh, operator, FSAI factor) resident in HBM under a
byte budget.  A repeated mesh reuses its system as is; a new mesh whose topology
is already cached refreshes a cached factor numerically on the stored supports
(``refit_clone``, the reuse path of fsai_reuse.py) instead of rerunning the
adaptive search; an unseen topology builds a factor.  Systems are built on a
worker thread (setup overlaps running solves) and live on a pool of HIP streams
(default priority); every admitted request owns a *slot* (its Krylov workspace,
captured as a HIP graph of ``chunk`` CG steps) on its system's stream.  The loop
admits arrived requests smallest estimated work first whenever a slot and budget
are free, and launches chunks for the jobs with the least remaining work first up
to a GPU-fill cap, so requests that arrive while others are on the GPU join at
chunk boundaries and their work overlaps on the device; host reads of solver state
are asynchronous (pinned buffers + events).

Stopping is certified: a request finishes when, for every right-hand side, an
upper bound on the relative mass-weighted field error is <= tol.  The float64
iterate is improved by iterative refinement: each outer step solves A c = r/||r||
with preconditioned CG to a relative residual ``inner_tol`` and adds the scaled
correction.  The bound is the smaller of refinement.field_bounds(x') =
||b - A x'||_{M^-1} / (alpha ||x'||_M) and refinement.witness_bounds, which uses
the correction itself: ||(r - A c)||_{M^-1} / (alpha ||x'||_M) + 2 eps, all
residuals recomputed with compensated float64.  The witness form matters for the
factored biharmonic operator, whose lambda_max(M^-1 A)/alpha reaches ~1e10 at
96^3: there the rounding of x' to float64 alone puts ~1e-7 into the first form,
while the true error is ~1e-16 of that size.  Two inner precisions:

* ``"f64"``   -- float64 CG on the float64 operator with a float32 factor;
* ``"mixed"`` -- float32 CG on a float32 copy of the operator with a row-scaled
  bfloat16 factor (built in float64).

Results do not depend on what else is in flight: every kernel reduces in a fixed
order, row mappings are chosen without timing, and a converged right-hand side's
further steps are no-ops, so a request's solution is bit-identical alone and in
any mix given the same factor provenance (built / cached / refit).
"""

from __future__ import annotations

import ctypes
import time
from collections import OrderedDict, deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native as nat
from . import assembly, grouped
from .fsai import FSAI
from .fsai_reuse import refit_clone
from .linear import KrylovWork
from .refinement import TrueResidual, field_bounds, witness_bounds
from .sparse import BsrMatrix
from .sparse_operator import SparseOperator
from .squared_laplacian import SquaredLaplacianOperator

KINDS = ("laplacian", "biharmonic")
_ITER, _RR, _TARGET = 4, 1, 2  # Krylov state fields (include/wpc.h WPC_KS_*)
# tightest inner relative residual worth asking of each inner precision
_INNER_FLOOR = {"f64": 1e-12, "mixed": 1e-6}


@dataclass(frozen=True)
class MeshSpec:
    side: int  # vertices per edge of the unit cube
    seed: int  # jitter seed


@dataclass
class Request:
    rid: int
    arrival: float  # seconds after the stream starts
    mesh: MeshSpec
    kind: str
    nrhs: int
    rhs_seed: int
    alpha: float = 1.0


def request_stream(count=512, rate=200.0, seed=7, sides=(16, 96), repeat=0.5, max_rhs=8,
                   kinds=KINDS):
    """Poisson arrivals (mean `rate`/s); each request is, with probability `repeat`,
    a uniformly chosen earlier (mesh, kind) system, otherwise a new mesh of side
    U{sides} with a uniformly chosen kind; 1..max_rhs right-hand sides."""
    rng = np.random.default_rng(seed)
    t = 0.0
    seen, out = [], []
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        if seen and rng.random() < repeat:
            mesh, kind = seen[int(rng.integers(len(seen)))]
        else:
            mesh = MeshSpec(int(rng.integers(sides[0], sides[1] + 1)), int(rng.integers(1 << 31)))
            kind = kinds[int(rng.integers(len(kinds)))]
            seen.append((mesh, kind))
        out.append(Request(rid, t, mesh, kind, int(rng.integers(1, max_rhs + 1)),
                           int(rng.integers(1 << 31))))
    return out


@dataclass
class ServingConfig:
    budget_bytes: float = 200e9
    streams: int = 4
    max_active: int = 64
    # FSAI width: an int, or {kind: int | [(smallest side, width), ...]} (the width of
    # the last entry whose side the mesh reaches, e.g. [(0, 8), (64, 16)])
    width: object = 8
    tol: float = 1e-8
    precision: dict = field(default_factory=lambda: {"laplacian": "mixed", "biharmonic": "f64"})
    factor_dtype: dict = field(default_factory=lambda: {"mixed": torch.bfloat16,
                                                        "f64": torch.float32})
    inner_tol: dict = field(default_factory=lambda: {"mixed": 1e-4, "f64": 1e-6})
    chunk: int = 25
    maxiter: int = 200000
    max_outer: int = 16
    keep_solutions: bool = True
    # one right-hand side per job, for the one-solve-at-a-time comparison path
    split_rhs: bool = False
    # advance the CG of the solves of one kind and format together
    # (grouped.GroupedCG: one launch per pass for the whole batch).  Off by
    # default: on the 64-request stream (sides 16-64) batches averaged 2.1-2.4
    # jobs, tied small solves to their partners' step time and took them off
    # their own streams (14.7-27.0 vs 27.2-28.1 solves/s); it is meant for many
    # same-size small solves in flight at once.
    group: bool = False


@dataclass
class Result:
    rid: int
    kind: str
    side: int
    n: int
    nrhs: int
    arrival: float
    admitted: float
    finished: float
    iterations: list
    outer: int
    field_bound: list
    relative_residual: list
    converged: bool
    factor: str  # "built", "refit" or "cached"
    x: np.ndarray | None = None
    history: list = field(default_factory=list)  # per outer step: (iterations, bounds)

    @property
    def latency(self):
        return self.finished - self.arrival


def _convert(t: torch.Tensor, dtype) -> torch.Tensor:
    out = torch.empty(t.shape, dtype=dtype, device=t.device)
    nat.call("wpc_convert", t.numel(), nat.ptr(t), nat.code(t.dtype), nat.ptr(out), nat.code(dtype),
             nat.stream())
    return out


class System:
    """Device-resident mesh operator and FSAI factor (built on the setup stream)."""

    def __init__(self, cfg: ServingConfig, kind, spec: MeshSpec, alpha, stream, template=None):
        if kind not in KINDS:
            raise ValueError(f"unknown system kind {kind!r}")
        dev = torch.device("cuda", torch.cuda.current_device())
        self.kind, self.spec, self.alpha, self.stream = kind, spec, float(alpha), stream
        self.policy = cfg.precision[kind]
        self.group = cfg.group
        self.slots = {}
        self.active = 0
        self.bytes = 0
        vertices, tets = assembly.jittered_grid(spec.side, spec.seed, dev)
        K, mass = assembly.p1_laplacian(vertices, tets)
        del vertices, tets
        n = K.nrow
        self.n, self.mass = n, mass
        ident = torch.arange(n, dtype=torch.int32, device=dev)
        if kind == "laplacian":
            A = assembly.add_diagonal(K, mass, self.alpha)
            del K
            self.op64 = SparseOperator(A, row_lanes=1)
            self.op64.source = self.op64.scalar
            del A
            source = self.op64.scalar
        else:
            self.op64 = SquaredLaplacianOperator(K, mass, ident, row_lanes=1,
                                                 mass_weight=self.alpha)
            del K
            source = assembly.squared_operator(self.op64.L, mass, self.alpha)
        if template is not None:
            self.P = refit_clone(template.P, source, topology=template.topology)
            self.topology = template.topology
            self.provenance = "refit"
        else:
            width = fsai_width(cfg.width, kind, spec.side)
            self.P = FSAI(source, max_row_size=width, factor_dtype=cfg.factor_dtype[self.policy])
            self.topology = (source.offsets, source.columns)
            self.provenance = "built"
        self.P.release_source()
        del source
        if self.policy == "mixed":
            if kind == "laplacian":
                S = self.op64.scalar
                self.inner = SparseOperator(
                    BsrMatrix(n, n, (1, 1), S.offsets, S.columns, _convert(S.values, torch.float32)),
                    row_lanes=1)
            else:
                L = self.op64.L
                self.inner = SquaredLaplacianOperator(
                    BsrMatrix(n, n, (1, 1), L.offsets, L.columns, _convert(L.values, torch.float32)),
                    _convert(mass, torch.float32), ident, row_lanes=1, mass_weight=self.alpha)
            self.vdtype = torch.float32
        else:
            self.inner = self.op64
            self.vdtype = torch.float64
        # Row mappings chosen without timing (results independent of load), and
        # every scratch buffer sized for 8 right-hand sides up front, so captured
        # graphs never see a reallocation.
        plans = list(self.P.plans()) + [self.inner._stream()]
        for p in plans:
            if p.fixed is None:
                p.fixed = p.heuristic()
            p.partials(8)
            if p.fixed == p.SLICED:
                p._sliced()  # build the sliced copy here, not on the scheduling thread
        self.P._scratch(self.vdtype, n * 8)
        if isinstance(self.inner, SquaredLaplacianOperator):
            self.inner._scratch(self.vdtype, n * 8)
            self.inner.energy_partials(8)  # the fused CG passes' curvature partials

    def capture_stream(self):
        """A stream that only records graphs (never runs work): the capturing
        thread's private stream (a torch stream could be another system's or the
        setup stream, torch hands them out from a shared pool)."""
        return nat.capture_stream()

    def slot(self, nvec):
        for s in self.slots.setdefault(nvec, []):
            if not s.busy:
                return s
        s = Slot(self, nvec)
        self.slots[nvec].append(s)
        return s

    def release(self):
        for slots in self.slots.values():
            for s in slots:
                s.release()
        self.slots = {}


class Slot:
    """One request's device workspace on a system (reused by later requests)."""

    def __init__(self, system: System, nvec: int):
        self.system, self.nvec = system, nvec
        n = system.n
        dev = system.mass.device
        with torch.cuda.stream(system.stream):
            size = n * nvec
            self.X = torch.zeros(size, dtype=torch.float64, device=dev)
            self.B = torch.empty(size, dtype=torch.float64, device=dev)
            self.res = TrueResidual(system.op64, system.mass, nvec)
            vdt = system.vdtype
            self.Rin = torch.empty(size, dtype=vdt, device=dev)
            self.C = torch.empty(size, dtype=vdt, device=dev)
            self.c64 = torch.empty(size, dtype=torch.float64, device=dev)
            self.norms_w = torch.empty(2 * nvec, dtype=torch.float64, device=dev)
            self.work = KrylovWork("cg", system.inner, system.P, n, nvec, vdt, dev)
            # grouped launches keep the separate update step; a work that may join
            # them takes the same steps alone
            self.work.fuse_update = not system.group
            self.x_iter = self.C
            self.gkey = grouped.group_key(self.work) if system.group else None
            gs = grouped.scratch(self.work) if self.gkey is not None else None
        r, w = self.res, self.work
        self.bytes = tensor_bytes([self.X, self.B, self.Rin, self.C, self.c64, self.norms_w, r.r,
                                   r.t_hi, r.t_lo, r.partials, r.norms, w.r, w.p, w.ap, w.z,
                                   w.map, w.az, w.parts, w.state, w.dot_parts, gs])
        # one per chunk in flight (two at most): a chunk's state is read while the
        # next chunk's copy may be landing
        self.state_hosts = [torch.empty(12 * nvec, dtype=torch.float64, pin_memory=True)
                            for _ in range(2)]
        self.norms_host = torch.empty(2 * nvec, dtype=torch.float64, pin_memory=True)
        self.norms_w_host = torch.empty(2 * nvec, dtype=torch.float64, pin_memory=True)
        self.x_host = torch.empty(n * nvec, dtype=torch.float64, pin_memory=True)
        self.graph = None
        self.graph_steps = 0
        self.warm = False
        self.busy = False

    def chunk(self, steps):
        """`steps` CG steps on the current stream: eager the first time (loads
        code objects), then captured once into a HIP graph and replayed."""
        if self.graph is not None and self.graph_steps == steps:
            nat.call("wpc_graph_launch", self.graph, nat.stream())
            return
        if not self.warm:
            self.work.steps(self.x_iter, steps)
            self.warm = True
            return
        # recorded on a private stream (no pending cross-stream waits), launched here
        handle = ctypes.c_void_p()
        launch_stream = nat.stream()
        recorded = True
        with torch.cuda.stream(self.system.capture_stream()):
            nat.call("wpc_graph_begin", nat.stream())
            try:
                self.work.steps(self.x_iter, steps)
            except nat.NativeError:
                recorded = False
            finally:
                rc = nat.lib().wpc_graph_end(nat.stream(), ctypes.addressof(handle))
        if not recorded or rc != 0:
            # invalidated from outside: nothing recorded ran; run this chunk directly
            if rc == 0 and handle.value:
                nat.lib().wpc_graph_destroy(handle.value)
            self.work.steps(self.x_iter, steps)
            return
        if self.graph is not None:
            nat.lib().wpc_graph_destroy(self.graph)
        self.graph, self.graph_steps = handle.value, steps
        nat.call("wpc_graph_launch", self.graph, launch_stream)

    def release(self):
        if self.graph is not None:
            nat.lib().wpc_graph_destroy(self.graph)
            self.graph = None


class Job:
    """State machine of one admitted request (all device work on the system stream)."""

    def __init__(self, eng: "ServingEngine", req: Request, system: System, slot: Slot,
                 provenance: str, t_admit: float):
        self.eng, self.req, self.system, self.slot = eng, req, system, slot
        self.cfg = eng.cfg
        self.provenance = provenance
        self.t_admit = t_admit
        self.pending = deque()  # (event, what, host buffer of a state read)
        self.last_state = None  # host copy of the state after the last chunk read
        self.chunks_out = 0
        self.phase = "init"
        self.outer = 0
        self.mask = (1 << req.nrhs) - 1
        self.iterations = np.zeros(req.nrhs, dtype=np.int64)
        self.inner_iters = np.zeros(req.nrhs, dtype=np.int64)
        self.bnorm = None
        self.bounds = np.full(req.nrhs, np.inf)
        self.rel = np.ones(req.nrhs)
        self.far = True
        self.best = np.full(req.nrhs, np.inf)
        self.stalls = 0
        self.history = []
        self.chunks_launched = 0
        self.fill = gpu_fill(system.n, req.nrhs)
        self.want_check = False
        self.wants_group = False
        self.finished = False
        self.result = None
        slot.busy = True
        system.active += 1

    # ------------------------------------------------------------------ helpers
    def _record(self, what, host=None):
        ev = torch.cuda.Event()
        ev.record(self.system.stream)
        self.pending.append((ev, what, host))

    def _start_solve(self, b, x, tol):
        x.zero_()
        self.slot.work.start(b, x, tol, 0.0)

    def _residual(self):
        s = self.slot
        s.res.run(s.X, s.B)
        s.norms_host.copy_(s.res.norms, non_blocking=True)
        self._record("residual")

    def _finish(self):
        s = self.slot
        s.x_host[: s.X.numel()].copy_(s.X, non_blocking=True)
        self._record("result")
        self.phase = "result"

    # ------------------------------------------------------------------ driving
    def poll(self) -> bool:
        """Consume completed host reads in order."""
        did = False
        while self.pending and self.pending[0][0].query():
            _, what, host = self.pending.popleft()
            self._on(what, host)
            did = True
        return did

    def enqueue(self) -> bool:
        return not self.finished and self._enqueue()

    def remaining(self) -> float:
        """Estimated seconds of GPU work left (scheduling priority)."""
        est = estimated_iterations(self.req.kind, self.system.spec.side)
        done = self.chunks_launched * self.cfg.chunk
        left = max(est - done, self.cfg.chunk)
        return 1.0e-10 * left * self.system.n * (4 + self.req.nrhs)

    def _enqueue(self) -> bool:
        s = self.slot
        st = self.system.stream
        with torch.cuda.stream(st):
            if self.phase == "init":
                gen = torch.Generator(device=s.X.device)
                gen.manual_seed(self.req.rhs_seed)
                torch.randn((self.system.n, self.req.nrhs), generator=gen,
                            out=s.B.view(self.system.n, self.req.nrhs))
                s.X.zero_()
                self._residual()  # x = 0: r = b gives ||b||_{M^-1}
                self.phase = "wait_residual"
                return True
            if self.phase == "solve":
                if self.want_check:
                    if self.chunks_out == 0:
                        self._after_solve()
                        return True
                    return False
                depth = 2 if self.far else 1
                if self.chunks_out < depth:
                    if s.gkey is not None:
                        self.wants_group = True  # the engine launches it with its batch
                        return False
                    self._own_chunk()
                    return True
        return False

    def _own_chunk(self):
        s = self.slot
        s.chunk(self.cfg.chunk)
        host = s.state_hosts[self.chunks_launched % 2]
        host.copy_(s.work.state, non_blocking=True)
        self._record("state", host)
        self.chunks_out += 1
        self.chunks_launched += 1

    def launch_own_chunk(self):
        """This job's chunk with its own launches on its system stream."""
        with torch.cuda.stream(self.system.stream):
            self._own_chunk()

    def grouped_chunk_launched(self, stream):
        """Bookkeeping after the engine launched this job's chunk in a batch on
        `stream` (the state read is ordered after it on the same stream)."""
        s = self.slot
        host = s.state_hosts[self.chunks_launched % 2]
        with torch.cuda.stream(stream):
            host.copy_(s.work.state, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.pending.append((ev, "state", host))
        self.chunks_out += 1
        self.chunks_launched += 1

    def _after_solve(self):
        """Recursive targets met: enqueue the certification step."""
        s = self.slot
        self.want_check = False
        iters = self.last_state.view(12, -1)[_ITER].numpy()[: self.req.nrhs]
        self.iterations += iters.astype(np.int64)
        nat.call("wpc_refine_update", self.system.n, self.req.nrhs, nat.ptr(s.X), nat.ptr(s.C),
                 nat.code(s.C.dtype), nat.ptr(s.res.norms), self.mask, nat.ptr(s.c64), nat.stream())
        # witness certificate of the updated iterate: r - A c, in place over r
        s.res.run(s.c64, s.res.r, norms=s.norms_w)
        s.norms_w_host.copy_(s.norms_w, non_blocking=True)
        self._residual()
        self.phase = "wait_residual"

    def _on(self, what, host=None):
        s = self.slot
        k = self.req.nrhs
        if what == "state":
            self.chunks_out -= 1
            self.last_state = host
            st = host.view(12, -1).numpy()[:, :k]
            rr, target = st[_RR], st[_TARGET]
            self.far = bool(np.all(rr > 1e4 * target))
            if np.all(rr <= target) or np.max(self.iterations + st[_ITER]) >= self.cfg.maxiter:
                self.want_check = True
            return
        if what == "residual":
            norms = s.norms_host.numpy()[: 2 * k].copy()
            if self.bnorm is None:
                self.bnorm = np.sqrt(norms[:k])
                self.prev_norms = norms
                self._begin_solve(first=True, norms=norms)
                return
            witness = s.norms_w_host.numpy()[: 2 * k].copy()
            plain = field_bounds(norms, k, self.system.alpha)
            wit = witness_bounds(norms, witness, self.prev_norms, k, self.system.alpha)
            fresh = np.minimum(plain, wit)
            # a right-hand side outside this step's mask kept its iterate, so the
            # certificate it already had still holds
            moved = np.array([(self.mask >> v) & 1 for v in range(k)], dtype=bool)
            self.bounds = np.where(moved, fresh, np.minimum(self.bounds, fresh))
            self.history.append((self.iterations.tolist(), plain.tolist(), wit.tolist()))
            self.prev_norms = norms
            self.rel = np.sqrt(norms[:k]) / np.where(self.bnorm > 0, self.bnorm, 1.0)
            self._begin_solve(first=False, norms=norms)
            return
        if what == "result":
            self.finished = True
            eng = self.eng
            x = None
            if self.cfg.keep_solutions:
                x = s.x_host[: self.system.n * k].numpy().reshape(self.system.n, k).copy()
            self.result = Result(
                self.req.rid, self.req.kind, self.system.spec.side, self.system.n, k,
                self.req.arrival, self.t_admit, eng.clock(), [int(v) for v in self.iterations],
                self.outer, [float(v) for v in self.bounds], [float(v) for v in self.rel],
                bool(np.all(self.bounds <= self.cfg.tol)), self.provenance, x, self.history)

    def _begin_solve(self, first, norms):
        """After a residual: finish, or start/continue the Krylov phase."""
        k = self.req.nrhs
        tol = self.cfg.tol
        s = self.slot
        if first:
            self.rel = np.ones(k)
            fail = self.bnorm > 0
        else:
            fail = self.bounds > tol
        if not np.any(fail):
            with torch.cuda.stream(self.system.stream):
                self._finish()
            return
        if not first:
            improved = self.bounds < 0.5 * self.best
            self.best = np.minimum(self.best, self.bounds)
            self.stalls = 0 if np.any(improved & fail) else self.stalls + 1
            total = int(np.max(self.iterations))
            if self.stalls >= 3 or total >= self.cfg.maxiter or self.outer >= self.cfg.max_outer:
                with torch.cuda.stream(self.system.stream):
                    self._finish()
                return
        self.outer += 1
        self.far = True
        self.mask = int(sum(1 << v for v in range(k) if fail[v]))
        policy = self.system.policy
        inner = self.cfg.inner_tol[policy]
        if not first:
            # aim the correction at what is still missing (one more step, usually)
            need = 0.1 * tol / float(np.max(self.bounds[fail]))
            inner = min(inner, max(_INNER_FLOOR[policy], need))
        with torch.cuda.stream(self.system.stream):
            nat.call("wpc_refine_prepare", self.system.n, k, nat.ptr(s.res.r), nat.ptr(s.res.norms),
                     self.mask, nat.ptr(s.Rin), nat.code(s.Rin.dtype), nat.stream())
            self._start_solve(s.Rin, s.C, inner)
        self.phase = "solve"

    def release(self):
        self.slot.busy = False
        self.system.active -= 1


def tensor_bytes(*objs) -> int:
    """Bytes of the distinct device storages reachable from the given objects'
    attributes (used for budget accounting: exact under concurrent allocation)."""
    seen_obj, seen_store = set(), set()
    total = 0
    stack = list(objs)
    while stack:
        o = stack.pop()
        if o is None or id(o) in seen_obj:
            continue
        seen_obj.add(id(o))
        if isinstance(o, torch.Tensor):
            if o.is_cuda:
                st = o.untyped_storage()
                if st.data_ptr() not in seen_store:
                    seen_store.add(st.data_ptr())
                    total += st.nbytes()
            continue
        if isinstance(o, (list, tuple)):
            stack.extend(o)
        elif isinstance(o, dict):
            stack.extend(o.values())
        elif hasattr(o, "__dict__") and type(o).__module__.startswith(__package__):
            stack.extend(v for k, v in vars(o).items() if k not in ("system", "eng", "template"))
    return total


def estimated_iterations(kind: str, side: int) -> float:
    """Inner CG steps per right-hand side to a certified 1e-8 on the jittered grid of
    `side` (fitted to measurements: sides 16..96, refinement included)."""
    return 6.3 * side if kind == "laplacian" else 1.6 * side * side


def estimated_seconds(req: Request, cached: bool) -> float:
    n = req.mesh.side ** 3
    work = 1.0e-10 * estimated_iterations(req.kind, req.mesh.side) * n * (4 + req.nrhs)
    return work + (0.0 if cached else 0.2 * n / 96 ** 3)


def fsai_width(width, kind: str, side: int) -> int:
    """The FSAI width of a system of `kind` on a mesh of `side` under the config's
    width policy (ServingConfig.width)."""
    w = width[kind] if isinstance(width, dict) else width
    if isinstance(w, int):
        return w
    chosen = None
    for smallest, value in sorted(w):
        if side >= smallest:
            chosen = value
    if chosen is None:
        raise ValueError(f"no FSAI width for side {side} in {w}")
    return int(chosen)


class PrioritySetup:
    """One setup worker thread that takes the pending job with the smallest
    priority first (ties in submission order), so a small request's system is not
    built behind a large one's."""

    def __init__(self):
        import heapq
        import itertools
        import threading

        self._heapq = heapq
        self._heap = []
        self._seq = itertools.count()
        self._cv = threading.Condition()
        self._closed = False
        self._thread = threading.Thread(target=self._run, name="wpc-setup", daemon=True)
        self._thread.start()

    def submit(self, priority, fn, *args):
        from concurrent.futures import Future

        fut = Future()
        with self._cv:
            if self._closed:
                raise RuntimeError("setup worker is shut down")
            self._heapq.heappush(self._heap, (priority, next(self._seq), fut, fn, args))
            self._cv.notify()
        return fut

    def _run(self):
        while True:
            with self._cv:
                while not self._heap and not self._closed:
                    self._cv.wait()
                if not self._heap:
                    return
                _, _, fut, fn, args = self._heapq.heappop(self._heap)
            if not fut.set_running_or_notify_cancel():
                continue
            try:
                fut.set_result(fn(*args))
            except BaseException as exc:  # handed to the scheduling thread
                fut.set_exception(exc)

    def shutdown(self, wait=True):
        with self._cv:
            self._closed = True
            self._cv.notify_all()
        if wait:
            self._thread.join()


def gpu_fill(n: int, nrhs: int) -> float:
    """Fraction of the device one system's products keep busy (~1 from about two
    million unknown-RHS pairs up)."""
    return min(1.0, n * nrhs / 2.0e6)


class ServingEngine:
    """Admission, factor cache and the chunk scheduler (one GPU; scheduling on the
    calling thread, system setup on one worker thread)."""

    FILL_CAP = 1.5  # launches go to the highest-priority jobs up to this much GPU fill
    AGING = 0.5  # admission priority: estimated seconds - AGING * seconds waited

    def __init__(self, config: ServingConfig | None = None, device=None):
        import threading

        self.cfg = config or ServingConfig()
        self.device = torch.device("cuda", 0) if device is None else torch.device(device)
        torch.cuda.set_device(self.device)
        self.streams = [torch.cuda.Stream(self.device) for _ in range(self.cfg.streams)]
        self.setup_stream = torch.cuda.Stream(self.device)
        self._next = 0
        self.systems = OrderedDict()  # (kind, mesh, alpha) -> System, LRU order
        self.templates = {}  # (kind, side, alpha, policy) -> System whose factor is refit
        self._templates_lock = threading.Lock()
        self._pool = PrioritySetup()
        self.building = {}  # key -> (future, estimated bytes)
        self.waiting = {}  # key -> [(request, admitted at)]
        self.t0 = None
        self.groups = {}  # group key -> (GroupedCG, stream)
        self.stats = dict(built=0, refit=0, cached=0, evicted=0, over_budget=0, setup_s=0.0,
                          peak_bytes=0, grouped_chunks=0, grouped_jobs=0)

    def clock(self):
        return time.perf_counter() - self.t0

    def __del__(self):
        pool = getattr(self, "_pool", None)
        if pool is not None:
            pool.shutdown(wait=False)  # its thread ends once no setup is pending

    def used_bytes(self):
        return (sum(s.bytes + sum(x.bytes for v in s.slots.values() for x in v)
                    for s in self.systems.values())
                + sum(b for _, b in self.building.values()))

    @staticmethod
    def estimate(req: Request, policy: str):
        """Upper bounds of a system's bytes (measured 900 B per unknown for Laplacian+mass,
        1,210 for the biharmonic kind, whose build also holds the assembled Q, ~780) and
        of a slot's (exact: 56 or 96 B per unknown and right-hand side)."""
        n = req.mesh.side ** 3
        system = (2000 if req.kind == "biharmonic" else 1000) * n
        slot = (96 if policy == "f64" else 56) * n * req.nrhs
        return system, slot

    def _make_room(self, need, keep=None):
        while self.used_bytes() + need > self.cfg.budget_bytes:
            victim = next((k for k, s in self.systems.items() if s.active == 0 and s is not keep),
                          None)
            if victim is None:
                return False
            s = self.systems.pop(victim)
            with self._templates_lock:
                for key, t in list(self.templates.items()):
                    if t is s:
                        del self.templates[key]
            s.release()
            self.stats["evicted"] += 1
        return True

    # ------------------------------------------------------------ setup worker
    def _build(self, req: Request, stream):
        """Runs on the setup thread: assemble the mesh operator, build or refit the
        factor on the setup stream; returns (system, ready event, seconds)."""
        torch.cuda.set_device(self.device)
        t = time.perf_counter()
        policy = self.cfg.precision[req.kind]
        tkey = (req.kind, req.mesh.side, req.alpha, policy)
        with self._templates_lock:
            template = self.templates.get(tkey)
        with torch.cuda.stream(self.setup_stream):
            s = System(self.cfg, req.kind, req.mesh, req.alpha, stream, template)
            ready = torch.cuda.Event()
            ready.record(self.setup_stream)
        s.bytes = tensor_bytes(s)
        if template is None:
            with self._templates_lock:
                self.templates.setdefault(tkey, s)
        return s, ready, time.perf_counter() - t

    def _collect_builds(self, active):
        for key, (fut, _) in list(self.building.items()):
            if not fut.done():
                continue
            del self.building[key]
            s, ready, seconds = fut.result()
            s.stream.wait_event(ready)
            self.stats["setup_s"] += seconds
            self.stats[s.provenance] += 1
            self.systems[key] = s
            for k, (req, t_admit) in enumerate(self.waiting.pop(key, [])):
                if k:
                    self.stats["cached"] += 1
                active.append(Job(self, req, s, s.slot(req.nrhs),
                                  s.provenance if k == 0 else "cached", t_admit))
            self.stats["peak_bytes"] = max(self.stats["peak_bytes"], self.used_bytes())

    # ------------------------------------------------------------ admission
    def _admit(self, req: Request, active, force=False) -> bool:
        key = (req.kind, req.mesh, req.alpha)
        s = self.systems.get(key)
        sys_b, slot_b = self.estimate(req, self.cfg.precision[req.kind])
        need = (0 if (s is not None or key in self.building) else sys_b) + slot_b
        if not self._make_room(need, keep=s):
            if not force:
                return False
            self.stats["over_budget"] += 1
        if s is not None:
            self.systems.move_to_end(key)
            self.stats["cached"] += 1
            active.append(Job(self, req, s, s.slot(req.nrhs), "cached", self.clock()))
        elif key in self.building:
            self.waiting[key].append((req, self.clock()))
        else:
            stream = self.streams[self._next % len(self.streams)]
            self._next += 1
            self.building[key] = (self._pool.submit(estimated_seconds(req, False), self._build,
                                                    req, stream), sys_b)
            self.waiting[key] = [(req, self.clock())]
        self.stats["peak_bytes"] = max(self.stats["peak_bytes"], self.used_bytes())
        return True

    def _priority(self, req: Request, now: float) -> float:
        cached = (req.kind, req.mesh, req.alpha) in self.systems
        return estimated_seconds(req, cached) - self.AGING * max(0.0, now - req.arrival)

    def serve(self, requests, realtime=True, on_result=None, deadline=None):
        """Run the request list to completion; returns Results in request order.
        With ``realtime`` each request becomes visible at its arrival time; otherwise
        all are available at once.  Arrived requests are admitted smallest estimated
        work first (aged by waiting time) while slots and budget allow; chunk launches
        go to the jobs with the least estimated remaining work first, up to FILL_CAP
        of GPU fill, so small requests run side by side and a large one alone.
        With ``deadline`` (seconds on the engine clock) the loop stops there: the
        unfinished requests' entries are None and ``self.timed_out`` is set."""
        self.t0 = time.perf_counter()
        self.timed_out = False
        if self.cfg.split_rhs:
            return self._serve_split(requests, realtime, on_result)
        queue = deque(sorted(requests, key=lambda r: r.arrival))
        ready, active, results = [], [], {}
        try:
            while queue or ready or active or self.building:
                now = self.clock()
                if deadline is not None and now > deadline:
                    self.timed_out = True
                    break
                while queue and (not realtime or queue[0].arrival <= now):
                    ready.append(queue.popleft())
                self._collect_builds(active)
                while ready:
                    pending = len(active) + sum(len(v) for v in self.waiting.values())
                    if pending >= self.cfg.max_active:
                        break
                    req = min(ready, key=lambda r: self._priority(r, now))
                    if not self._admit(req, active, force=pending == 0):
                        break
                    ready.remove(req)
                progressed = False
                for job in active:
                    progressed |= job.poll()
                fill = 0.0
                batches = {}
                for job in sorted(active, key=Job.remaining):
                    if fill >= self.FILL_CAP:
                        break
                    if not job.finished:
                        job.wants_group = False
                        progressed |= job.enqueue()
                        if job.wants_group:
                            batches.setdefault(job.slot.gkey, []).append(job)
                        fill += job.fill
                for key, jobs in batches.items():
                    if len(jobs) > 1:
                        self._launch_batch(key, jobs)
                    else:  # a batch of one keeps the job's own stream (concurrency)
                        jobs[0].launch_own_chunk()
                    progressed = True
                done = [j for j in active if j.finished]
                for j in done:
                    active.remove(j)
                    j.release()
                    results[j.req.rid] = j.result
                    if on_result is not None:
                        on_result(j.result)
                if not progressed and not done:
                    if active or self.building:
                        time.sleep(50e-6)
                    elif queue and realtime:
                        time.sleep(max(0.0, min(queue[0].arrival - self.clock(), 0.01)))
        finally:
            for fut, _ in self.building.values():
                fut.cancel()
        if self.timed_out:
            torch.cuda.synchronize(self.device)  # in-flight chunks end before teardown
        return [results.get(r.rid) for r in requests]

    def _launch_batch(self, key, jobs):
        """One chunk of CG steps for every job in `jobs` (same group key) with one
        launch per pass, on the batch's own stream, ordered after each job's
        earlier work and before its later work."""
        entry = self.groups.get(key)
        if entry is None:
            entry = (grouped.GroupedCG(key, self.cfg.chunk), torch.cuda.Stream(self.device))
            self.groups[key] = entry
        group, stream = entry
        for j in jobs:
            ev = torch.cuda.Event()
            ev.record(j.system.stream)
            stream.wait_event(ev)
        with torch.cuda.stream(stream):
            group.run([j.slot.work for j in jobs], [j.slot.x_iter for j in jobs])
        for j in jobs:
            j.grouped_chunk_launched(stream)
        done = torch.cuda.Event()
        done.record(stream)
        for j in jobs:
            j.system.stream.wait_event(done)
        self.stats["grouped_chunks"] += 1
        self.stats["grouped_jobs"] += len(jobs)

    def _serve_split(self, requests, realtime, on_result):
        """Every right-hand side of every request as a job of its own (nvec = 1);
        a request finishes with its last right-hand side."""
        parts, owner = [], {}
        for req in requests:
            for v in range(req.nrhs):
                sub = Request(len(parts), req.arrival, req.mesh, req.kind, 1,
                              req.rhs_seed * 8 + v, req.alpha)
                owner[sub.rid] = req
                parts.append(sub)
        self.cfg.split_rhs = False
        try:
            sub_results = self.serve(parts, realtime, None)
        finally:
            self.cfg.split_rhs = True
        merged = {}
        for sub, res in zip(parts, sub_results):
            req = owner[sub.rid]
            m = merged.get(req.rid)
            if m is None:
                merged[req.rid] = m = Result(req.rid, req.kind, res.side, res.n, 0, req.arrival,
                                             res.admitted, res.finished, [], 0, [], [], True,
                                             res.factor, None)
            m.nrhs += 1
            m.admitted = min(m.admitted, res.admitted)
            m.finished = max(m.finished, res.finished)
            m.iterations += res.iterations
            m.outer = max(m.outer, res.outer)
            m.field_bound += res.field_bound
            m.relative_residual += res.relative_residual
            m.converged = m.converged and res.converged
        out = [merged[r.rid] for r in requests]
        if on_result is not None:
            for r in out:
                on_result(r)
        return out
