"""One batched CG step over a table of systems (serving extension).

``GroupedCG`` advances the fused preconditioned CG of several ``KrylovWork``s --
systems of uneven size and right-hand-side count, one vector format -- with
one launch per pass for all of them (``wpc_sell_stream_grouped``,
``wpc_cg_fold_grouped``): the operator pass(es), the update, the two FSAI passes
and the direction step.  Each system runs exactly the workgroups its own launch
would run, so its iterates are bit-identical to ``KrylovWork.step``.

The launches read device tables of per-system pointers.  A chunk of steps is
captured once as a HIP graph sized for a capacity of systems and workgroups;
later chunks rewrite the tables (one host-to-device copy) and replay the graph,
so systems join and leave the batch without a recapture.  Unused capacity is
padded with empty tasks.  Applies to works whose products use the sliced mapping
and whose inner products fit the in-place folds (``group_key`` says which).
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _native as nat
from .fsai import FSAI
from .linear import FOLD_IN_PLACE, KrylovWork
from .sparse import StreamPlan
from .sparse_operator import SparseOperator
from .squared_laplacian import SquaredLaplacianOperator

_WORDS = 16  # 64-bit words per task (include/wpc.h, grouped CG steps)


def _p(t):
    """Device address of a tensor for a task table (0 for none)."""
    return 0 if t is None else t.data_ptr()


def _sell_blocks(n):
    return (n + 255) // 256


def _operator_kind(A):
    if isinstance(A, SquaredLaplacianOperator) and A.identity:
        return "factored"
    if isinstance(A, SparseOperator):
        return "csr"
    return None


def group_key(work: KrylovWork):
    """Hashable description of the launches a work needs, or None if it cannot
    join a grouped step (not fused CG, other operators, other row mappings, or
    partials too many for the in-place folds)."""
    if not work.use_fused or work.method != 0:
        return None
    A, M = work.A, work.M
    kind = _operator_kind(A)
    if kind is None or not isinstance(M, FSAI) or work.dtype not in (torch.float32, torch.float64):
        return None
    plans = [A._stream(), *M.plans()]
    if any(p.fixed != StreamPlan.SLICED for p in plans):
        return None
    n, nv = work.n, work.nvec
    if max(_sell_blocks(n), nat.lib().wpc_krylov_partials(n)) * nv > FOLD_IN_PLACE:
        return None
    a_vals = (A.L if kind == "factored" else A.scalar).values.dtype
    weight = kind == "factored" and A.mass_weight != 0.0
    return (kind, a_vals, weight, M.G.values.dtype, M.row_scale2 is not None, work.dtype)


class _Scratch:
    """Per-work buffers a grouped step needs of its own (the operators' scratch and
    partials are shared by every work on one system)."""

    def __init__(self, work: KrylovWork):
        size = work.n * work.nvec
        dev = work.r.device
        self.tmp_a = torch.empty(size, dtype=work.dtype, device=dev)
        self.tmp_m = torch.empty(size, dtype=work.dtype, device=dev)
        parts = _sell_blocks(work.n) * work.nvec
        self.part_ap = torch.zeros(parts, dtype=torch.float64, device=dev)
        self.part_rz = torch.zeros(parts, dtype=torch.float64, device=dev)


def scratch(work: KrylovWork) -> _Scratch:
    s = getattr(work, "_grouped_scratch", None)
    if s is None:
        s = work._grouped_scratch = _Scratch(work)
    return s


def _sliced(plan):
    sp, sl, sc, sv = plan._sliced()
    return [_p(sp), _p(sl), _p(plan._sell[5]), _p(sc), _p(sv)]


def _col16(plan):
    """The task's slice_base word (the per-slice column bases, or the row masks of
    the stencil form)."""
    return _p(plan._sell[7])


def _nv_word(plan, nv):
    """The task's vector-count word: nv, plus 256 times the column form (1: 16-bit
    offsets, 2: stencil)."""
    return nv | (256 * int(plan._sell[6]))


class GroupedCG:
    """Captured grouped CG chunks for one ``group_key``."""

    def __init__(self, key, steps):
        self.key = key
        self.steps = steps
        self.kind, self.a_dtype, self.weight, self.g_dtype, self.scaled, self.dtype = key
        self.graph = None
        self.cap_tasks = 0
        self.cap_sell = 0
        self.cap_fold = 0
        self.tables = None
        self._ring = [(None, None)] * 4
        self._ring_next = 0
        self.warm = False
        self.capture_failures = 0

    # ------------------------------------------------------------ tables
    def _passes(self, works):
        """Host tables (int64 [tasks, 14]) of every pass, in launch order."""
        ops = []
        sell = []
        fold_u, fold_d = [], []
        b_sell = b_fold = 0
        for w in works:
            sc = scratch(w)
            A, M = w.A, w.M
            n = w.n
            nb = nat.lib().wpc_krylov_partials(n)
            pa = _sell_blocks(n)
            plan_a = A._stream()
            p_, r_, z_, ap_ = (_p(t) for t in (w.p, w.r, w.z, w.ap))
            a_sl = _sliced(plan_a)
            pg, pgt = M.plans()
            ca, cg, cgt = _col16(plan_a), _col16(pg), _col16(pgt)
            nv = w.nvec
            if self.kind == "factored":
                k1 = a_sl + [_p(A.inverse_mass), 0, 0, p_, _p(sc.tmp_a), 0, 0, n, b_sell, ca,
                             _nv_word(plan_a, nv)]
                k2 = a_sl + [0, _p(A._wm) if self.weight else 0, p_ if self.weight else 0,
                             _p(sc.tmp_a), ap_, p_, _p(sc.part_ap), n, b_sell, ca,
                             _nv_word(plan_a, nv)]
                ops.append((k1, k2))
            else:
                ops.append((a_sl + [0, 0, 0, p_, ap_, p_, _p(sc.part_ap), n, b_sell, ca,
                                    _nv_word(plan_a, nv)],))
            g = _sliced(pg) + [_p(M.row_scale2) if self.scaled else 0, 0, 0, r_,
                               _p(sc.tmp_m), 0, 0, n, b_sell, cg, _nv_word(pg, nv)]
            gt = _sliced(pgt) + [0, 0, 0, _p(sc.tmp_m), z_, r_, _p(sc.part_rz), n,
                                 b_sell, cgt, _nv_word(pgt, nv)]
            sell.append((g, gt))
            pr = _p(w.parts[1])
            st = _p(w.state)
            # (the update leaves x to the direction step, which advances it before
            # replacing p)
            fold_u.append([0, r_, 0, ap_, 0, _p(sc.part_ap), 0, st, pr,
                           n, nb, b_fold, pa, 0, nv, 0])
            fold_d.append([_p(w._group_x), 0, p_, 0, z_, _p(sc.part_rz), pr, st, 0, n, nb, b_fold,
                           pa, nb, nv, 0])
            b_sell += pa
            b_fold += nb
        nops = len(ops[0])
        tables = [[o[i] for o in ops] for i in range(nops)]
        tables.append(fold_u)
        tables += [[s_[0] for s_ in sell], [s_[1] for s_ in sell]]
        tables.append(fold_d)
        return tables, b_sell, b_fold

    def _fill(self, tables, b_sell, b_fold):
        """Copy the member tables into the device tables (padding tasks at the end)."""
        hosts = []
        nfold = len(tables)
        for i, rows in enumerate(tables):
            is_fold = i == nfold - 1 or i == nfold - 4
            out = np.zeros((self.cap_tasks + 1, _WORDS), dtype=np.int64)
            out[: len(rows)] = np.array(rows, dtype=np.int64)
            # padding: n = 0 and block0 = the end of the real workgroups
            out[len(rows):, 11 if is_fold else 13] = b_fold if is_fold else b_sell
            if is_fold:
                out[len(rows):, 10] = 1  # nblocks
                out[len(rows):, 14] = 1  # nv
            else:
                out[len(rows):, 15] = 1  # nv
            hosts.append(out)
        # pinned staging, reused once its previous copy has completed
        k = self._ring_next
        self._ring_next = (k + 1) % len(self._ring)
        buf, ev = self._ring[k]
        if buf is None or buf.shape != self.tables.shape:
            buf = torch.empty(self.tables.shape, dtype=torch.int64, pin_memory=True)
            ev = torch.cuda.Event()
        else:
            ev.synchronize()
        buf.numpy()[...] = np.stack(hosts)
        self.tables.copy_(buf, non_blocking=True)
        ev.record()
        self._ring[k] = (buf, ev)

    # ------------------------------------------------------------ launches
    def _step(self):
        s = nat.stream()
        t = self.tables
        nt = self.cap_tasks + 1
        code = nat.code(self.dtype)
        a_code = nat.code(self.a_dtype)
        g_code = nat.code(self.g_dtype)
        sell = lambda i, vals, s_code, has_d, dot: nat.call(  # noqa: E731
            "wpc_sell_stream_grouped", nt, nat.ptr(t[i]), self.cap_sell, vals, s_code, has_d, dot,
            code, 1.0, s)
        fold = lambda step, i: nat.call(  # noqa: E731
            "wpc_cg_fold_grouped", step, nt, nat.ptr(t[i]), self.cap_fold, code, s)
        i = 0
        if self.kind == "factored":
            sell(0, a_code, code, 0, 0)  # K pass 1, inverse-mass scale
            sell(1, a_code, 0, int(self.weight), 1)  # K pass 2 (+ weight M p, p.Ap partials)
            i = 2
        else:
            sell(0, a_code, 0, 0, 1)  # A p (+ p.Ap partials)
            i = 1
        fold(1, i)  # update
        sell(i + 1, g_code, nat.F32 if self.scaled else 0, 0, 0)  # G pass
        sell(i + 2, g_code, 0, 0, 1)  # G^T pass (+ r.z partials)
        fold(2, i + 3)  # direction

    def _capture(self):
        """Record `steps` steps into a graph on this thread's private capture
        stream, which never runs work (a stream with pending cross-stream waits, or
        one another owner shares, is not captured cleanly), to be launched on the
        caller's stream."""
        if self.graph is not None:
            nat.lib().wpc_graph_destroy(self.graph)
            self.graph = None
        handle = ctypes.c_void_p()
        with torch.cuda.stream(nat.capture_stream()):
            nat.call("wpc_graph_begin", nat.stream())
            try:
                for _ in range(self.steps):
                    self._step()
            finally:
                rc = nat.lib().wpc_graph_end(nat.stream(), ctypes.addressof(handle))
        if rc != 0:
            raise nat.NativeError(f"wpc_graph_end failed: {rc}")
        self.graph = handle.value

    def run(self, works, xs, capture=True):
        """``steps`` CG steps for every work (x_i the iterate of works[i]) on the
        current stream.  Works must share this key and be distinct."""
        for w, x in zip(works, xs):
            w._group_x = x
        tables, b_sell, b_fold = self._passes(works)
        grow = (len(works) > self.cap_tasks or b_sell > self.cap_sell or b_fold > self.cap_fold)
        if grow or self.tables is None:
            self.cap_tasks = max(self.cap_tasks, int(len(works) * 1.25) + 1)
            self.cap_sell = max(self.cap_sell, int(b_sell * 1.25) + 1)
            self.cap_fold = max(self.cap_fold, int(b_fold * 1.25) + 1)
            dev = works[0].r.device
            self.tables = torch.zeros((len(tables), self.cap_tasks + 1, _WORDS), dtype=torch.int64,
                                      device=dev)
            if self.graph is not None:
                nat.lib().wpc_graph_destroy(self.graph)
                self.graph = None
        self._fill(tables, b_sell, b_fold)
        # (the null stream cannot be captured: chunks launched there stay eager)
        if not capture or not self.warm or nat.stream() in (0, None):
            # the first chunk runs eagerly (loads the code objects), later ones replay
            for _ in range(self.steps):
                self._step()
            self.warm = True
            return
        if self.graph is None:
            try:
                self._capture()
            except nat.NativeError:
                # A capture can be invalidated from outside (e.g. a synchronizing
                # call on another thread); nothing it recorded has run, so this
                # chunk runs eagerly and a later chunk captures again.
                self.graph = None
                self.capture_failures += 1
                for _ in range(self.steps):
                    self._step()
                return
        nat.call("wpc_graph_launch", self.graph, nat.stream())

    def release(self):
        if self.graph is not None:
            nat.lib().wpc_graph_destroy(self.graph)
            self.graph = None
