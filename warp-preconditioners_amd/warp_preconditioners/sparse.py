"""Block-sparse-row matrices over torch device buffers.

The reference package stores matrices in Warp's ``warp.sparse.BsrMatrix`` and uses
Warp's ``bsr_from_triplets``, ``bsr_copy``, ``bsr_zeros``, ``bsr_transposed`` and
``bsr_mv`` (call sites: fsai.py:279,325,337; sparse_operator.py:104,127;
biharmonic.py:97-106,135,169-170).  ``BsrMatrix`` here exposes the attributes
the package reads (offsets, columns, values, nrow, ncol, shape, block_shape,
scalar_type, dtype, device, nnz, nnz_sync(), row_counts) with int32 index
buffers and values of shape [nnz] (1x1 blocks) or [nnz, b, b].  Storage is
always compact and canonical: sorted columns, no duplicate blocks.
"""

from __future__ import annotations

import ctypes

from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat


@dataclass(frozen=True)
class BlockType:
    """dtype of a b x b block (Warp's matrix(shape=(b, b)) value type)."""

    rows: int
    cols: int
    scalar_type: torch.dtype


class BsrMatrix:
    def __init__(self, nrow, ncol, block_shape, offsets, columns, values):
        self.nrow = int(nrow)
        self.ncol = int(ncol)
        self.block_shape = tuple(block_shape)
        self.offsets = offsets
        self.columns = columns
        self.values = values
        self.row_counts = None  # always compact

    @property
    def scalar_type(self) -> torch.dtype:
        return self.values.dtype

    @property
    def dtype(self):
        if self.block_shape == (1, 1):
            return self.values.dtype
        return BlockType(self.block_shape[0], self.block_shape[1], self.values.dtype)

    @property
    def device(self) -> torch.device:
        return self.values.device

    @property
    def shape(self):
        return (self.nrow * self.block_shape[0], self.ncol * self.block_shape[1])

    @property
    def nnz(self) -> int:
        return int(self.columns.numel())

    def nnz_sync(self) -> int:
        return self.nnz

    def notify_nnz_changed(self, nnz=None):
        pass

    # host helpers (tests, diagnostics) ------------------------------------
    def to_scipy(self):
        import scipy.sparse as ss

        b0, b1 = self.block_shape
        off = self.offsets.cpu().numpy()
        col = self.columns.cpu().numpy()
        val = self.values.detach().to(torch.float64).cpu().numpy()
        if (b0, b1) == (1, 1):
            return ss.csr_matrix((val.reshape(-1), col, off), shape=self.shape)
        return ss.bsr_matrix((val.reshape(-1, b0, b1), col, off), shape=self.shape).tocsr()

    def to_dense(self) -> np.ndarray:
        return self.to_scipy().toarray()


class StreamPlan:
    """Launch plan for ``wpc_csr_stream`` on one scalar CSR matrix.

    Three row mappings exist: streamed (mode 0: row blocks of <= 256 rows / 2048
    entries staged through LDS, computed once on the host from the offsets),
    multi-lane (mode L: L lanes per row) and sliced (mode -1: a sliced-ELL copy of
    the matrix, one thread per row; bit-identical to mode 0).  The sliced copy is
    built on first use and refreshed by ``refresh_values`` after the CSR values
    change in place.  ``lanes=None`` autotunes: on the first
    product for a given (vector count, vector dtype) outside graph capture, every
    mode is timed with HIP events on the current stream and the fastest is kept;
    under capture a row-length heuristic decides.  The fused inner-product
    partials belong to the plan, so one plan must not run on two streams at once.
    """

    CANDIDATES = (0, 2, 4, 8, 16, -1)
    SLICED = -1
    # Choices are shared by every plan over the same index buffers in this process,
    # so re-planning one matrix never changes its summation order.
    _chosen = {}

    def __init__(self, A: "BsrMatrix", lanes=None):
        if A.block_shape != (1, 1):
            raise ValueError("streamed products need a scalar CSR matrix")
        off = np.ascontiguousarray(A.offsets.cpu().numpy(), dtype=np.int32)
        cap = A.nrow + 2
        blocks = np.empty(cap, dtype=np.int32)
        count = nat.lib().wpc_csr_row_blocks(A.nrow, off.ctypes.data, blocks.ctypes.data, cap)
        if count < 0:
            raise nat.NativeError("wpc_csr_row_blocks failed")
        self.nblocks = int(count)
        self.blocks = torch.from_numpy(blocks[: count + 1].copy()).to(A.device)
        self.fixed = lanes
        self.modes = {}
        self.timings = {}
        self._partials = None
        self._sell = None  # (slice_ptr, slice_len, columns, map)
        self._sell_values = None
        self.A = A
        self.mean_row = float(off[-1]) / max(A.nrow, 1)

    @classmethod
    def like(cls, other: "StreamPlan", A: "BsrMatrix") -> "StreamPlan":
        """Plan for a matrix with `other`'s index buffers (same row partition and
        row-mapping choices, no host work); owns its own inner-product partials."""
        if A.offsets.data_ptr() != other.A.offsets.data_ptr():
            raise ValueError("StreamPlan.like requires the same offsets buffer")
        plan = cls.__new__(cls)
        plan.nblocks, plan.blocks, plan.fixed = other.nblocks, other.blocks, other.fixed
        plan.modes, plan.timings = dict(other.modes), {}
        plan._partials = None
        plan._sell = other._sell
        plan._sell_values = None
        plan.A = A
        plan.mean_row = other.mean_row
        return plan

    # Rows are ordered by length inside windows of SORT_WINDOW rows only when that
    # removes more than SORT_GAIN * nnz padded entries.  Natural order keeps the
    # gathers of neighbouring rows together, which outweighs moderate padding: on
    # the 64^3 biharmonic G^T (28% of the stored entries are padding) the 3-vector
    # pass with its dot takes 9.7 us in natural order and 10.2 / 9.9 / 10.5 / 11.7 us
    # ordered in windows of 128 / 256 / 512 / 1024 rows (padding 18 / 11 / 6 / 4%).
    # The order pays for rows that are far longer than their slice neighbours
    # (padding above the matrix's own size).
    SORT_WINDOW = 256
    SORT_GAIN = 1.0
    # Store the sliced columns as int16 differences from the row when they all fit.
    COL16 = True
    # Column forms of the sliced copy (the products' col16 argument).
    COL_I32, COL_U16, COL_STENCIL = 0, 1, 2
    # Use the stencil form (wpc_sell_stencil: per-slice column offsets from the row
    # and per-row masks, no column stream) when every slice's rows share at most 32
    # offsets and the form stores at most 1/64 more entries than the plain slices
    # (mesh-ordered Laplacians: 15 offsets on a Kuhn grid).
    STENCIL = True

    def _row_order(self):
        """Row order of the sliced copy (device int32 permutation) or None."""
        A = self.A
        n = A.nrow
        if n < 128:
            return None
        ln = np.diff(A.offsets.cpu().numpy().astype(np.int64))
        nnz = int(ln.sum())

        def padded(lengths):
            m = np.zeros(-(-n // 64) * 64, np.int64)
            m[:n] = lengths
            return int(m.reshape(-1, 64).max(axis=1).sum()) * 64

        order = np.lexsort((-ln, np.arange(n) // self.SORT_WINDOW))  # stable
        if padded(ln) - padded(ln[order]) <= self.SORT_GAIN * max(nnz, 1):
            return None
        return torch.from_numpy(order.astype(np.int32)).to(A.device)

    def _sliced(self):
        """(slice_ptr, slice_len, columns, values) of the sliced copy, built on
        first use (host reads of the offsets and of the padded size)."""
        A = self.A
        if self._sell is None:
            ns = nat.lib().wpc_sell_slices(A.nrow)
            dev = A.device
            rows = self._row_order()
            slice_len = torch.empty(max(ns, 1), dtype=torch.int32, device=dev)
            slice_ptr = torch.zeros(ns + 1, dtype=torch.int32, device=dev)
            s = nat.stream()
            nat.call("wpc_sell_lengths", A.nrow, nat.ptr(A.offsets), nat.ptr(rows),
                     nat.ptr(slice_len), s)
            if ns:
                nat.call("wpc_scan_inclusive", ns, nat.ptr(slice_len), nat.ptr(slice_ptr[1:]), s)
            total = int(slice_ptr[-1].item()) * 64
            stencil = self._stencil(ns, total) if rows is None and total else None
            if stencil is not None:
                self._sell = stencil
                total = stencil[4]
            else:
                self._sell = self._plain(ns, rows, slice_ptr, slice_len, total)
        if self._sell_values is None:
            self._sell_values = torch.empty(max(self._sell[4], 1), dtype=A.values.dtype,
                                            device=A.device)
            self.refresh_values()
        return self._sell[0], self._sell[1], self._sell[2], self._sell_values

    def _stencil(self, ns, total):
        """The stencil form of the slices (wpc_sell_stencil) as a `_sell` tuple, or
        None where a slice has more than 32 offsets or the form would store more
        than 1/64 more entries than the `total` of the plain slices (on Kuhn grids
        whose side is not a multiple of 64 a few slices span grid lines whose rows
        miss different neighbours: 0.1-0.3% more entries; G^T's 30% stays plain)."""
        if not self.STENCIL:
            return None
        A = self.A
        dev, s = A.device, nat.stream()
        slice_len = torch.empty(ns, dtype=torch.int32, device=dev)
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        nat.call("wpc_sell_stencil", A.nrow, nat.ptr(A.offsets), nat.ptr(A.columns), None,
                 nat.ptr(slice_len), None, None, None, nat.ptr(status), s)
        slice_ptr = torch.zeros(ns + 1, dtype=torch.int32, device=dev)
        nat.call("wpc_scan_inclusive", ns, nat.ptr(slice_len), nat.ptr(slice_ptr[1:]), s)
        stotal = int(slice_ptr[-1].item()) * 64
        if int(status.item()) != 0 or stotal > total + total // 64:
            return None
        deltas = torch.empty(ns * 32, dtype=torch.int32, device=dev)
        masks = torch.empty(A.nrow, dtype=torch.int32, device=dev)
        smap = torch.empty(max(stotal, 1), dtype=torch.int32, device=dev)
        nat.call("wpc_sell_stencil", A.nrow, nat.ptr(A.offsets), nat.ptr(A.columns),
                 nat.ptr(slice_ptr), nat.ptr(slice_len), nat.ptr(deltas), nat.ptr(masks),
                 nat.ptr(smap), nat.ptr(status), s)
        # (the row masks travel in the slice_base argument)
        return (slice_ptr, slice_len, deltas, smap, stotal, None, self.COL_STENCIL, masks)

    def _plain(self, ns, rows, slice_ptr, slice_len, total):
        """The plain slices (int32 columns, or 16-bit offsets from per-slice bases
        when every slice's columns span at most 0xFFFE) as a `_sell` tuple."""
        A = self.A
        dev, s = A.device, nat.stream()
        cols = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        smap = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        nat.call("wpc_sell_fill", A.nrow, nat.ptr(A.offsets), nat.ptr(A.columns),
                 nat.ptr(rows), nat.ptr(slice_ptr), nat.ptr(slice_len), nat.ptr(cols),
                 nat.ptr(smap), s)
        col16 = self.COL_I32
        # per-slice column bases of the 16-bit copy (zeros with int32 columns)
        base = torch.zeros(max(ns, 1), dtype=torch.int32, device=dev)
        if self.COL16 and total:
            c16 = torch.empty(total, dtype=torch.int16, device=dev)
            over = torch.zeros(1, dtype=torch.int32, device=dev)
            nat.call("wpc_sell_col16", A.nrow, nat.ptr(A.offsets), nat.ptr(rows),
                     nat.ptr(slice_ptr), nat.ptr(slice_len), nat.ptr(cols), nat.ptr(c16),
                     nat.ptr(base), nat.ptr(over), s)
            if int(over.item()) == 0:
                cols, col16 = c16, self.COL_U16
            else:
                base.zero_()  # int32 columns: the descriptor starts at x
        return (slice_ptr, slice_len, cols, smap, total, rows, col16, base)

    # The sliced product gathers int32 columns through a buffer descriptor whose
    # byte offsets (column * nvec * element size) must stay below 2^31.
    DESCRIPTOR_RANGE = 1 << 31

    def _sell_fits(self, x, nvec):
        return self.sliced_fits(nvec, x.element_size())

    def sliced_fits(self, nvec, elem_bytes):
        if self._sell is not None and self._sell[6] == self.COL_U16:
            return True  # 16-bit offsets: the range is per slice
        return self.A.ncol * nvec * elem_bytes <= self.DESCRIPTOR_RANGE

    def refresh_values(self):
        """Re-gather the sliced copy's values after A.values changed in place."""
        if self._sell is not None and self._sell_values is not None:
            A = self.A
            nat.call("wpc_sell_values", self._sell[4], nat.ptr(self._sell[3]), nat.ptr(A.values),
                     A.values.element_size(), nat.ptr(self._sell_values), nat.stream())

    def heuristic(self):
        """Row mapping chosen without timing: the sliced mapping, or 8 lanes per row
        for rows too long to pad (its copy is built on first use, so under capture
        only an already-built copy qualifies)."""
        if self.mean_row > 64:
            return 8
        if torch.cuda.is_current_stream_capturing() and self._sell_values is None:
            return 0  # same sums as the sliced mapping, no copy to build
        return self.SLICED

    def workgroups(self, mode):
        if mode == self.SLICED:
            return (self.A.nrow + 255) // 256
        return self.nblocks if mode == 0 else (self.A.nrow * mode + 255) // 256

    def partials_size(self, nvec, mode=None):
        """Doubles the fused inner-product partials need: one per workgroup and
        vector of the widest mapping this plan can launch (every tuning candidate,
        the fixed lane count, and `mode`)."""
        modes = set(self.CANDIDATES)
        if self.fixed is not None:
            modes.add(self.fixed)
        if mode is not None:
            modes.add(mode)
        return max(max(self.workgroups(m) for m in modes), 1) * nvec

    def partials(self, nvec, mode=None):
        need = self.partials_size(nvec, mode)
        if self._partials is None or self._partials.numel() < need:
            self._partials = torch.zeros(need, dtype=torch.float64, device=self.A.device)
            nat.bump_generation()
        return self._partials

    def mode(self, nvec, dtype):
        if self.fixed is not None:
            return self.fixed
        key = (nvec, dtype)
        m = self.modes.get(key)
        if m is None:
            A = self.A
            shared = (A.offsets.data_ptr(), A.columns.data_ptr(), A.nrow, A.nnz, nvec, dtype)
            m = StreamPlan._chosen.get(shared)
            if m is None:
                if torch.cuda.is_current_stream_capturing():
                    return self.heuristic()
                m = self._tune(nvec, dtype)
                StreamPlan._chosen[shared] = m
            self.modes[key] = m
        return m

    # A mapping must beat the sliced one (storage-order sums, bit-identical to the
    # streamed mapping and the serial oracle) by this margin to be chosen instead.
    SLICED_MARGIN = 0.03

    def _tune(self, nvec, dtype, reps=10, rounds=3):
        """Time every candidate mapping; candidates are interleaved over `rounds`
        rounds and each keeps its best round, so a transient stall on the device
        does not decide the choice.  Each candidate's `reps` launches are captured
        into a graph (the library's thread-local capture: no device-wide sync) and
        replayed, as the solvers replay their iterations, so the host's launch cost
        (which can exceed a small product's run time) does not hide the difference
        between mappings; direct launches are timed if a capture fails."""
        n = self.A.nrow
        x = torch.ones(n * nvec, dtype=dtype, device=self.A.device)
        z = torch.empty_like(x)
        work = torch.cuda.Stream()
        work.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(work):
            for m in self.CANDIDATES:
                for _ in range(2):
                    self._launch(m, x, None, z, 1.0, 0.0, nvec)
        work.synchronize()
        # captured on this thread's private stream, which runs nothing
        graphs = {}
        with torch.cuda.stream(nat.capture_stream()):
            for m in self.CANDIDATES:
                if nat.lib().wpc_graph_begin(nat.stream()) != 0:
                    break
                recorded = True
                handle = ctypes.c_void_p()
                try:
                    for _ in range(reps):
                        self._launch(m, x, None, z, 1.0, 0.0, nvec)
                except nat.NativeError:
                    recorded = False  # (the capture is still ended below)
                finally:
                    rc = nat.lib().wpc_graph_end(nat.stream(), ctypes.addressof(handle))
                if rc != 0 or not recorded:
                    if rc == 0 and handle.value:
                        nat.lib().wpc_graph_destroy(handle.value)
                    break
                graphs[m] = handle.value
        if len(graphs) < len(self.CANDIDATES):
            for g in graphs.values():
                nat.lib().wpc_graph_destroy(g)
            graphs = {}
        try:
            with torch.cuda.stream(work):
                for _ in range(rounds):
                    for m in self.CANDIDATES:
                        e0 = torch.cuda.Event(enable_timing=True)
                        e1 = torch.cuda.Event(enable_timing=True)
                        e0.record(work)
                        if graphs:
                            nat.call("wpc_graph_launch", graphs[m], nat.stream())
                        else:
                            for _ in range(reps):
                                self._launch(m, x, None, z, 1.0, 0.0, nvec)
                        e1.record(work)
                        e1.synchronize()
                        t = e0.elapsed_time(e1) / reps
                        key = (nvec, dtype, m)
                        self.timings[key] = min(t, self.timings.get(key, float("inf")))
        finally:
            for g in graphs.values():
                nat.lib().wpc_graph_destroy(g)
        torch.cuda.current_stream().wait_stream(work)
        best, best_t = 0, float("inf")
        for m in self.CANDIDATES:
            t = self.timings[(nvec, dtype, m)]
            if t < best_t:
                best, best_t = m, t
        sliced = self.timings.get((nvec, dtype, self.SLICED))
        if sliced is not None and self.mean_row <= 64 and sliced <= best_t * (1 + self.SLICED_MARGIN):
            return self.SLICED
        return best

    def _launch(self, mode, x, y, z, alpha, beta, nvec, s=None, d=None, u=None, w=None, out=None,
                values=None, part=None):
        A = self.A
        if mode == self.SLICED and values is None and not self._sell_fits(x, nvec):
            mode = 0  # same sums, no descriptor range limit
        if mode == self.SLICED and values is None:
            sp, sl, sc, sv = self._sliced()
            nat.call("wpc_sell_stream", A.nrow, nat.ptr(sp), nat.ptr(sl), nat.ptr(self._sell[7]),
                     nat.ptr(self._sell[5]), nat.ptr(sc), int(self._sell[6]), nat.ptr(sv),
                     nat.code(sv.dtype), nat.ptr(s), nat.code(s.dtype) if s is not None else 0,
                     nat.ptr(d), nat.ptr(u), nat.ptr(x), nat.ptr(y), nat.ptr(z), nat.code(x.dtype),
                     nvec, float(alpha), float(beta), nat.ptr(w), nat.ptr(part), nat.ptr(out),
                     nat.stream())
            return
        if mode == self.SLICED:
            mode = 0  # explicit values: the streamed mapping sums in the same order
        vals = A.values if values is None else values
        nat.call("wpc_csr_stream", self.nblocks, nat.ptr(self.blocks), A.nrow, mode,
                 nat.ptr(A.offsets), nat.ptr(A.columns), nat.ptr(vals), nat.code(vals.dtype),
                 nat.ptr(s), nat.code(s.dtype) if s is not None else 0, nat.ptr(d), nat.ptr(u),
                 nat.ptr(x), nat.ptr(y), nat.ptr(z), nat.code(x.dtype), nvec, float(alpha),
                 float(beta), nat.ptr(w), nat.ptr(part), nat.ptr(out), nat.stream())

    def run(self, x, y, z, alpha, beta, nvec=1, s=None, d=None, u=None, w=None, out=None,
            values=None):
        """z = alpha*(s .* (A x) + d .* u) + beta*y.

        With ``w``: per-workgroup partials of w . z are left in the plan's buffer
        (and folded into ``out`` when given); returns (partials, count)."""
        mode = self.mode(nvec, x.dtype)
        part = self.partials(nvec, mode) if w is not None else None
        self._launch(mode, x, y, z, alpha, beta, nvec, s, d, u, w, out, values, part)
        return part, self.workgroups(mode)


def time_graph(launch, reps=10, rounds=3):
    """Best time (ms) of one `launch()` over `rounds` rounds of `reps` launches
    replayed from a HIP graph captured on this thread's private stream (direct
    launches are timed if capture fails); host launch cost does not enter."""
    work = torch.cuda.Stream()
    work.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(work):
        launch()
        launch()
    work.synchronize()
    graph = None
    with torch.cuda.stream(nat.capture_stream()):
        if nat.lib().wpc_graph_begin(nat.stream()) == 0:
            handle = ctypes.c_void_p()
            ok = True
            try:
                for _ in range(reps):
                    launch()
            except nat.NativeError:
                ok = False
            finally:
                rc = nat.lib().wpc_graph_end(nat.stream(), ctypes.addressof(handle))
            if rc == 0 and handle.value:
                if ok:
                    graph = handle.value
                else:
                    nat.lib().wpc_graph_destroy(handle.value)
    best = float("inf")
    try:
        with torch.cuda.stream(work):
            for _ in range(rounds):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(work)
                if graph is not None:
                    nat.call("wpc_graph_launch", graph, nat.stream())
                else:
                    for _ in range(reps):
                        launch()
                e1.record(work)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1) / reps)
    finally:
        if graph is not None:
            nat.lib().wpc_graph_destroy(graph)
    torch.cuda.current_stream().wait_stream(work)
    return best


class BsrPlan:
    """Sliced copy of a 3x3 BSR matrix and its product (``wpc_bsr3_sell_stream``).

    Block rows in slices of 64 (one wavefront), block columns stored as the scalar
    sliced copy stores columns (built by the same ``wpc_sell_*`` kernels on the
    block pattern; 16-bit offsets from per-slice bases when they fit), the nine
    values of each block component-major inside the slice.  One thread per block
    row; storage-order sums, bit-identical to the oracle's ``bsr_spmv``.  The
    fused inner-product partials belong to the plan (one stream at a time).
    """

    COL16 = True

    def __init__(self, A: "BsrMatrix"):
        if A.block_shape != (3, 3):
            raise ValueError("BsrPlan needs 3x3 blocks")
        if A.scalar_type not in (torch.float32, torch.float64):
            raise TypeError("BsrPlan supports float32 and float64 values")
        self.A = A
        nb = A.nrow
        dev = A.device
        ns = nat.lib().wpc_sell_slices(nb)
        s = nat.stream()
        self.slice_len = torch.empty(max(ns, 1), dtype=torch.int32, device=dev)
        self.slice_ptr = torch.zeros(ns + 1, dtype=torch.int32, device=dev)
        nat.call("wpc_sell_lengths", nb, nat.ptr(A.offsets), None, nat.ptr(self.slice_len), s)
        if ns:
            nat.call("wpc_scan_inclusive", ns, nat.ptr(self.slice_len), nat.ptr(self.slice_ptr[1:]), s)
        self.slots = int(self.slice_ptr[-1].item()) * 64
        cols = torch.empty(max(self.slots, 1), dtype=torch.int32, device=dev)
        self.map = torch.empty(max(self.slots, 1), dtype=torch.int32, device=dev)
        nat.call("wpc_sell_fill", nb, nat.ptr(A.offsets), nat.ptr(A.columns), None,
                 nat.ptr(self.slice_ptr), nat.ptr(self.slice_len), nat.ptr(cols), nat.ptr(self.map), s)
        self.slice_base = torch.zeros(max(ns, 1), dtype=torch.int32, device=dev)
        self.col16 = False
        if self.COL16 and self.slots:
            c16 = torch.empty(self.slots, dtype=torch.int16, device=dev)
            over = torch.zeros(1, dtype=torch.int32, device=dev)
            nat.call("wpc_sell_col16", nb, nat.ptr(A.offsets), None, nat.ptr(self.slice_ptr),
                     nat.ptr(self.slice_len), nat.ptr(cols), nat.ptr(c16), nat.ptr(self.slice_base),
                     nat.ptr(over), s)
            if int(over.item()) == 0:
                cols, self.col16 = c16, True
            else:
                self.slice_base.zero_()
        self.columns = cols
        self.values = torch.empty(max(self.slots, 1) * 9, dtype=A.scalar_type, device=dev)
        self.refresh_values()
        self._partials = None

    def refresh_values(self):
        """Re-gather the sliced values after A.values changed in place."""
        nat.call("wpc_bsr3_sell_values", self.slots, nat.ptr(self.map), nat.ptr(self.A.values),
                 self.A.values.element_size(), nat.ptr(self.values), nat.stream())

    def workgroups(self):
        return (self.A.nrow + 255) // 256

    def partials(self, nvec):
        need = max(self.workgroups(), 1) * nvec
        if self._partials is None or self._partials.numel() < need:
            self._partials = torch.zeros(need, dtype=torch.float64, device=self.A.device)
            nat.bump_generation()
        return self._partials

    def fits(self, x, nvec):
        """The int32 columns' descriptor range (ncol * 3 * nvec * element size <= 2^31)."""
        return self.col16 or self.A.ncol * 3 * nvec * x.element_size() <= StreamPlan.DESCRIPTOR_RANGE

    def run(self, x, y, z, alpha, beta, nvec=1, w=None, out=None):
        """z = alpha A x + beta y on nvec interleaved vectors; with w, per-workgroup
        partials of w . z (folded into `out` when given); returns (partials, count)."""
        A = self.A
        part = self.partials(nvec) if w is not None else None
        nat.call("wpc_bsr3_sell_stream", A.nrow, nat.ptr(self.slice_ptr), nat.ptr(self.slice_len),
                 nat.ptr(self.slice_base), nat.ptr(self.columns), int(self.col16),
                 nat.ptr(self.values), nat.code(A.scalar_type), nat.ptr(x), nat.ptr(y), nat.ptr(z),
                 nat.code(x.dtype), nvec, float(alpha), float(beta), nat.ptr(w), nat.ptr(part),
                 nat.ptr(out), nat.stream())
        return part, self.workgroups()


def _device(device):
    return torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())


def bsr_zeros(nrow, ncol, dtype, device=None, block_shape=(1, 1)) -> BsrMatrix:
    device = _device(device)
    shape = (0,) if tuple(block_shape) == (1, 1) else (0, *block_shape)
    return BsrMatrix(
        nrow, ncol, block_shape,
        torch.zeros(nrow + 1, dtype=torch.int32, device=device),
        torch.zeros(0, dtype=torch.int32, device=device),
        torch.zeros(shape, dtype=dtype, device=device),
    )


def bsr_from_triplets(nrow, ncol, rows, cols, values, prune_numerical_zeros=False) -> BsrMatrix:
    """Sum duplicate (row, col) blocks and sort (Warp's bsr_from_triplets).

    ``values`` is [nnz] for scalars or [nnz, b, b] for blocks.  Stored zeros are
    kept unless ``prune_numerical_zeros``.
    """
    rows = torch.as_tensor(rows).to(torch.int64)
    cols = torch.as_tensor(cols).to(torch.int64)
    values = torch.as_tensor(values)
    device = values.device
    rows, cols = rows.to(device), cols.to(device)
    block_shape = (1, 1) if values.dim() == 1 else tuple(values.shape[1:])
    key = rows * int(ncol) + cols
    order = torch.argsort(key, stable=True)
    key = key[order]
    uniq, inverse = torch.unique_consecutive(key, return_inverse=True)
    summed = torch.zeros((uniq.numel(), *values.shape[1:]), dtype=values.dtype, device=device)
    summed.index_add_(0, inverse, values[order])
    if prune_numerical_zeros:
        keep = summed.reshape(uniq.numel(), -1).ne(0).any(dim=1)
        uniq, summed = uniq[keep], summed[keep]
    r = uniq // int(ncol)
    c = (uniq % int(ncol)).to(torch.int32)
    counts = torch.bincount(r, minlength=int(nrow))
    offsets = torch.zeros(int(nrow) + 1, dtype=torch.int32, device=device)
    offsets[1:] = torch.cumsum(counts, 0).to(torch.int32)
    return BsrMatrix(nrow, ncol, block_shape, offsets, c.contiguous(), summed.contiguous())


def from_scipy(a, device=None, dtype=torch.float64, block_shape=(1, 1)) -> BsrMatrix:
    """Upload a scipy sparse matrix (canonical BSR with the given block shape)."""
    import scipy.sparse as ss

    device = _device(device)
    b0, b1 = block_shape
    if (b0, b1) == (1, 1):
        m = ss.csr_matrix(a, copy=True)
        m.sum_duplicates()
        m.sort_indices()
        vals = np.asarray(m.data)
        nrow, ncol = m.shape
    else:
        m = ss.bsr_matrix(a, blocksize=(b0, b1))
        m.sum_duplicates()
        m.sort_indices()
        vals = np.asarray(m.data)
        nrow, ncol = m.shape[0] // b0, m.shape[1] // b1
    np_dtype = {torch.float16: np.float16, torch.float32: np.float32, torch.float64: np.float64}[dtype]
    return BsrMatrix(
        nrow, ncol, (b0, b1),
        torch.from_numpy(np.ascontiguousarray(m.indptr, dtype=np.int32)).to(device),
        torch.from_numpy(np.ascontiguousarray(m.indices, dtype=np.int32)).to(device),
        torch.from_numpy(np.ascontiguousarray(vals.astype(np_dtype))).to(device),
    )


def scalarize(A: BsrMatrix) -> BsrMatrix:
    """Expand square b x b blocks into scalar CSR on the device, keeping stored zeros
    (fsai_pack.py:75-130)."""
    b = A.block_shape[0]
    if A.block_shape != (b, b):
        raise ValueError("scalarize requires square blocks")
    nat.require_device(A.values, "matrix")
    n = A.nrow * b
    count = A.nnz * b * b
    dev = A.device
    off = torch.empty(n + 1, dtype=torch.int32, device=dev)
    col = torch.empty(count, dtype=torch.int32, device=dev)
    val = torch.empty(count, dtype=A.scalar_type, device=dev)
    nat.call("wpc_bsr_scalarize", A.nrow, b, nat.ptr(A.offsets), nat.ptr(A.columns),
             nat.ptr(A.values), nat.code(A.scalar_type), nat.ptr(off), nat.ptr(col), nat.ptr(val),
             nat.stream())
    return BsrMatrix(n, A.ncol * b, (1, 1), off, col, val)


def bsr_copy(A: BsrMatrix, block_shape=None, scalar_type=None) -> BsrMatrix:
    """Copy with optional reblocking to 1x1 and/or value conversion (Warp's bsr_copy)."""
    out = A
    if block_shape is not None and tuple(block_shape) != A.block_shape:
        if tuple(block_shape) != (1, 1):
            raise NotImplementedError("bsr_copy reblocks to scalar (1, 1) blocks only")
        out = scalarize(A)
    else:
        out = BsrMatrix(A.nrow, A.ncol, A.block_shape, A.offsets.clone(), A.columns.clone(),
                        A.values.clone())
    if scalar_type is not None and scalar_type != out.scalar_type:
        vals = torch.empty(out.values.shape, dtype=scalar_type, device=out.device)
        nat.call("wpc_convert", out.values.numel(), nat.ptr(out.values), nat.code(out.scalar_type),
                 nat.ptr(vals), nat.code(scalar_type), nat.stream())
        out = BsrMatrix(out.nrow, out.ncol, out.block_shape, out.offsets, out.columns, vals)
    return out


def bsr_transposed(A: BsrMatrix) -> BsrMatrix:
    """Transpose of a scalar CSR matrix with sorted rows (Warp's bsr_transposed)."""
    if A.block_shape != (1, 1):
        raise NotImplementedError("bsr_transposed supports scalar matrices")
    dev = A.device
    nnz = A.nnz
    toff = torch.empty(A.ncol + 1, dtype=torch.int32, device=dev)
    tcol = torch.empty(nnz, dtype=torch.int32, device=dev)
    tval = torch.empty(nnz, dtype=A.scalar_type, device=dev)
    work = torch.empty(A.ncol + 1, dtype=torch.int32, device=dev)
    nat.call("wpc_csr_transpose", A.nrow, A.ncol, nnz, nat.ptr(A.offsets), nat.ptr(A.columns),
             nat.ptr(A.values), A.values.element_size(), nat.ptr(toff), nat.ptr(tcol),
             nat.ptr(tval), nat.ptr(work), nat.stream())
    return BsrMatrix(A.ncol, A.nrow, (1, 1), toff, tcol, tval)


def bsr_mv(A: BsrMatrix, x, y, alpha=1.0, beta=0.0, lanes=8):
    """y = alpha * A x + beta * y (Warp's bsr_mv semantics; y updated in place)."""
    xf = x.reshape(-1)
    yf = y.reshape(-1)
    if xf.data_ptr() == yf.data_ptr():
        xf = xf.clone()
    _product(A, xf, yf, yf, alpha, beta, lanes)
    return y


def _product(A: BsrMatrix, x, y, z, alpha, beta, lanes=8):
    """z = alpha A x + beta y on flat vectors (x must not alias z)."""
    b = A.block_shape[0]
    if A.block_shape == (1, 1):
        nat.call("wpc_csr_spmv", A.nrow, nat.ptr(A.offsets), nat.ptr(A.columns), nat.ptr(A.values),
                 nat.code(A.scalar_type), None, nat.ptr(x), nat.ptr(y), nat.ptr(z),
                 nat.code(x.dtype), 1, float(alpha), float(beta), lanes, nat.stream())
    else:
        if x.dtype != A.scalar_type:
            raise TypeError("block products require vectors of the matrix scalar type")
        nat.call("wpc_bsr_spmv", A.nrow, b, nat.ptr(A.offsets), nat.ptr(A.columns),
                 nat.ptr(A.values), nat.code(A.scalar_type), nat.ptr(x), nat.ptr(y), nat.ptr(z),
                 float(alpha), float(beta), nat.stream())
