"""numpy pipelines over the C++ oracle -- TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Each pipeline mirrors the reference's host logic step by step and raises the same
ValueError texts, so a parity test can compare the HIP product against it on the
same seeded inputs.
"""

from __future__ import annotations

import ctypes
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "liboracle.so"

F16, BF16, F32, F64 = 1, 2, 3, 4
_CODES = {np.dtype(np.float16): F16, np.dtype(np.float32): F32, np.dtype(np.float64): F64}

_lib = None


def build() -> None:
    """Compile the oracle restatement (g++, no fused multiply-add)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def code(dtype) -> int:
    return _CODES[np.dtype(dtype)]


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


@dataclass
class CSR:
    """Canonical compact BSR/CSR on the host: sorted columns, b x b blocks."""

    off: np.ndarray
    col: np.ndarray
    val: np.ndarray  # shape (nnzb,) for b == 1, (nnzb, b, b) otherwise
    nrow: int  # block rows
    ncol: int
    block: int = 1

    @property
    def nnz(self) -> int:
        return int(self.off[-1]) if len(self.off) else 0

    @property
    def dtype(self):
        return self.val.dtype

    @staticmethod
    def from_scipy(a, dtype=np.float64, block: int = 1) -> "CSR":
        import scipy.sparse as ss

        a = ss.csr_matrix(a)
        if block == 1:
            a = a.copy()
            a.sum_duplicates()
            a.sort_indices()
            return CSR(
                _i32(a.indptr), _i32(a.indices), np.ascontiguousarray(a.data, dtype=dtype),
                a.shape[0], a.shape[1], 1,
            )
        b = ss.bsr_matrix(a, blocksize=(block, block))
        b.sum_duplicates()
        b.sort_indices()
        return CSR(
            _i32(b.indptr), _i32(b.indices), np.ascontiguousarray(b.data, dtype=dtype),
            b.shape[0] // block, b.shape[1] // block, block,
        )

    def to_dense(self) -> np.ndarray:
        b = self.block
        out = np.zeros((self.nrow * b, self.ncol * b), dtype=np.float64)
        vals = self.val.astype(np.float64).reshape(-1, b, b)
        for i in range(self.nrow):
            for e in range(self.off[i], self.off[i + 1]):
                j = self.col[e]
                out[i * b:(i + 1) * b, j * b:(j + 1) * b] += vals[e]
        return out

    def to_scipy(self):
        import scipy.sparse as ss

        if self.block == 1:
            return ss.csr_matrix(
                (self.val.astype(np.float64), self.col, self.off), shape=(self.nrow, self.ncol)
            )
        return ss.bsr_matrix(
            (self.val.astype(np.float64), self.col, self.off),
            shape=(self.nrow * self.block, self.ncol * self.block),
        ).tocsr()


def scalarize(A: CSR) -> CSR:
    """fsai_pack.py:75-130 -- expand compact b x b BSR keeping stored zeros."""
    if A.block == 1:
        return A
    b = A.block
    n = A.nrow * b
    count = A.nnz * b * b
    off = np.zeros(n + 1, np.int32)
    col = np.zeros(count, np.int32)
    val = np.zeros(count, A.dtype)
    src = np.ascontiguousarray(A.val.reshape(-1))
    lib().orc_bsr_scalarize(code(A.dtype), A.nrow, b, _p(A.off), _p(A.col), _p(src),
                            _p(off), _p(col), _p(val))
    return CSR(off, col, val, n, A.ncol * b, 1)


def transpose(A: CSR) -> CSR:
    """Warp bsr_transposed (fsai.py:337) for scalar CSR."""
    toff = np.zeros(A.ncol + 1, np.int32)
    tcol = np.zeros(A.nnz, np.int32)
    tval = np.zeros(A.nnz, A.dtype)
    lib().orc_csr_transpose(A.dtype.itemsize, A.nrow, A.ncol, _p(A.off), _p(A.col), _p(A.val),
                            _p(toff), _p(tcol), _p(tval))
    return CSR(toff, tcol, tval, A.ncol, A.nrow, 1)


DEFAULT_FLOOR = {np.dtype(np.float16): 1.0e-3, np.dtype(np.float32): 1.0e-6,
                 np.dtype(np.float64): 1.0e-12}


def fsai_raw(S: CSR, width: int, step: int, tolerance: float, pivot_floor: float, nthreads=1):
    """fsai.py:284-312 -- diagonal + build on a scalar matrix; padded output."""
    n = S.nrow
    dt = code(S.dtype)
    scale = np.zeros(n, S.dtype)
    status = np.zeros(3, np.int32)
    lib().orc_fsai_diagonal(dt, n, _p(S.off), _p(S.col), _p(S.val), _p(scale), _p(status))
    cols = np.zeros(n * width, np.int32)
    vals = np.zeros(n * width, S.dtype)
    lib().orc_fsai_build(dt, n, _p(S.off), _p(S.col), _p(S.val), _p(scale),
                         ctypes.c_double(tolerance), ctypes.c_double(pivot_floor), width, step,
                         _p(cols), _p(vals), _p(status), nthreads)
    return scale, cols, vals, status


def pack_factor(n: int, width: int, cols, vals) -> CSR:
    """fsai_pack.py:77-90."""
    cols = cols.copy()
    vals = vals.copy()
    counts = np.zeros(n + 1, np.int32)
    lib().orc_fsai_sort_rows(code(vals.dtype), n, width, _p(cols), _p(vals), _p(counts))
    off = np.zeros(n + 1, np.int32)
    lib().orc_scan_inclusive(n + 1, _p(counts), _p(off))
    nnz = int(off[-1])
    outc = np.zeros(nnz, np.int32)
    outv = np.zeros(nnz, vals.dtype)
    lib().orc_fsai_pack(code(vals.dtype), n, width, _p(off), _p(cols), _p(vals), _p(outc),
                        _p(outv))
    return CSR(off, outc, outv, n, n, 1)


def convert(values: np.ndarray, dtype) -> np.ndarray:
    out = np.zeros(values.shape, dtype)
    lib().orc_convert(code(values.dtype), code(dtype), ctypes.c_long(values.size),
                      _p(np.ascontiguousarray(values)), _p(out))
    return out


def fsai(A: CSR, max_row_size=8, kap_tolerance=1.0e-3, pivot_floor=None, factor_dtype=None,
         max_step_size=1, nthreads=1):
    """fsai.py:222-344 -- returns dict(G, GT, truncated_rows, scalar, scale)."""
    if pivot_floor is None:
        pivot_floor = DEFAULT_FLOOR[np.dtype(A.dtype)]
    S = scalarize(A)
    n = S.nrow
    scale, cols, vals, status = fsai_raw(S, max_row_size, max_step_size, kap_tolerance,
                                         pivot_floor, nthreads)
    if status[0]:
        raise ValueError(f"FSAI requires finite positive diagonals ({status[0]} invalid rows)")
    truncated = int(status[1])
    if status[2]:
        raise ValueError("FSAI factor lost finite entries or a positive diagonal")
    G = pack_factor(n, max_row_size, cols, vals)
    if factor_dtype is not None and np.dtype(factor_dtype) != np.dtype(S.dtype):
        G = CSR(G.off, G.col, convert(G.val, factor_dtype), n, n, 1)
        st = np.zeros(1, np.int32)
        lib().orc_validate_factor(code(G.dtype), n, _p(G.off), _p(G.col), _p(G.val), _p(st))
        if st[0]:
            raise ValueError("factor_dtype conversion lost finite entries or a positive diagonal")
    GT = transpose(G)
    return dict(G=G, GT=GT, truncated_rows=truncated, scalar=S, scale=scale,
                pivot_floor=pivot_floor, width=max_row_size)


def block_fsai(A: CSR, max_blocks=2, factor_dtype=None, refit_from=None):
    """benchmarks/block_fsai_candidate.py:174-215 (and the refit of :217-254 when
    ``refit_from`` is a previous result): dict(G, GT, cols, vals, scale)."""
    b = A.block
    if b < 2:
        raise ValueError("block FSAI needs blocks of size 2 or more")
    nb = A.nrow
    n = nb * b
    width = b * max_blocks
    dt = code(A.dtype)
    scale = np.zeros(n, A.dtype)
    status = np.zeros(1, np.int32)
    if refit_from is not None:
        cols = refit_from["cols"].copy()
        vals = np.zeros(n * width, A.dtype)
    else:
        cols = np.zeros(n * width, np.int32)
        vals = np.zeros(n * width, A.dtype)
    val = np.ascontiguousarray(A.val.reshape(-1))
    lib().orc_block_fsai_build(dt, nb, b, max_blocks, _p(A.off), _p(A.col), _p(val), _p(scale),
                               _p(cols), _p(vals), _p(status), int(refit_from is not None))
    if status[0]:
        raise ValueError("Nonpositive block FSAI pivot")
    G = pack_factor(n, width, cols, vals)
    # fsai_pack.py:15-39 sorts the padded rows in place; a later refit reads its
    # block pattern from those sorted rows (block_fsai_candidate.py:76-82)
    counts = np.zeros(n + 1, np.int32)
    lib().orc_fsai_sort_rows(code(vals.dtype), n, width, _p(cols), _p(vals), _p(counts))
    if factor_dtype is not None and np.dtype(factor_dtype) != np.dtype(A.dtype):
        G = CSR(G.off, G.col, convert(G.val, factor_dtype), n, n, 1)
    return dict(G=G, GT=transpose(G), cols=cols, vals=vals, scale=scale, width=width)


def spmv(A: CSR, x, y=None, alpha=1.0, beta=0.0, vec_dtype=None):
    """sparse_operator.py:64-87 semantics (scalar CSR, accumulation in vector type)."""
    vec_dtype = np.dtype(vec_dtype or x.dtype)
    x = np.ascontiguousarray(x, vec_dtype)
    y = np.zeros(A.nrow, vec_dtype) if y is None else np.ascontiguousarray(y, vec_dtype)
    z = np.zeros(A.nrow, vec_dtype)
    lib().orc_csr_spmv(code(A.dtype), code(vec_dtype), A.nrow, _p(A.off), _p(A.col), _p(A.val),
                       _p(x), _p(y), _p(z), ctypes.c_double(alpha), ctypes.c_double(beta))
    return z


def fsai_apply(P, x, y=None, alpha=1.0, beta=0.0):
    """fsai.py:374-381 -- z = alpha G^T G x + beta y."""
    G, GT = P["G"], P["GT"]
    vd = np.dtype(x.dtype)
    x = np.ascontiguousarray(x)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, vd)
    tmp = np.zeros(G.nrow, vd)
    z = np.zeros_like(x)
    lib().orc_fsai_apply(code(G.dtype), code(vd), G.nrow, _p(G.off), _p(G.col), _p(G.val),
                         _p(GT.off), _p(GT.col), _p(GT.val), _p(tmp), _p(x), _p(y), _p(z),
                         ctypes.c_double(alpha), ctypes.c_double(beta))
    return z


def refit(P, A: CSR):
    """fsai_reuse.py:195-245 -- fixed-support numerical refit; returns (G, GT) values."""
    S_old = P["scalar"]
    G, GT = P["G"], P["GT"]
    S = scalarize(A)
    if not (np.array_equal(S.off, S_old.off) and np.array_equal(S.col, S_old.col)):
        raise ValueError("FSAI update requires identical sparsity; rebuild to change it")
    n = S.nrow
    dt = code(S.dtype)
    scale = np.zeros(n, S.dtype)
    bad = np.zeros(3, np.int32)
    lib().orc_fsai_diagonal(dt, n, _p(S.off), _p(S.col), _p(S.val), _p(scale), _p(bad))
    out = np.zeros(G.nnz, G.dtype)
    bad2 = np.zeros(1, np.int32)
    lib().orc_fsai_refit(dt, code(G.dtype), n, _p(S.off), _p(S.col), _p(S.val), _p(G.off),
                         _p(G.col), _p(scale), ctypes.c_double(P["pivot_floor"]), _p(out),
                         _p(bad2))
    if bad[0] or bad2[0]:
        raise ValueError("FSAI refit requires finite positive diagonals and SPD local supports")
    tmap = np.zeros(G.nnz, np.int32)
    lib().orc_refit_transpose_map(n, _p(G.off), _p(G.col), _p(GT.off), _p(GT.col), _p(tmap))
    tval = np.zeros(G.nnz, G.dtype)
    lib().orc_scatter(G.dtype.itemsize, ctypes.c_long(G.nnz), _p(out), _p(tmap), _p(tval))
    newG = CSR(G.off, G.col, out, n, n, 1)
    newGT = CSR(GT.off, GT.col, tval, n, n, 1)
    return dict(P, G=newG, GT=newGT, scalar=S, scale=scale)


def source_map(S: CSR, A: CSR) -> np.ndarray:
    """fsai_reuse.py:35-48."""
    mapping = np.zeros(S.nnz, np.int32)
    lib().orc_refit_source_map(S.nrow, _p(S.off), _p(S.col), _p(A.off), _p(A.col), A.block,
                               _p(mapping))
    return mapping


# ---------------------------------------------------------------- squared Laplacian
def sqlap_setup(L: CSR, mass: np.ndarray, free: np.ndarray):
    """squared_laplacian.py:120-152 (validation and inverse mass)."""
    n = L.nrow
    index = np.zeros(n, np.int32)
    status = np.zeros(3, np.int32)
    free = _i32(free)
    lib().orc_sqlap_index_free(len(free), _p(free), n, _p(index), _p(status))
    inverse = np.zeros(n, L.dtype)
    mass = np.ascontiguousarray(mass, L.dtype)
    lib().orc_sqlap_mass_inverse(code(L.dtype), n, _p(mass), _p(inverse), _p(status))
    if status[0] or status[1]:
        raise ValueError("free_indices must be distinct and in bounds")
    if status[2]:
        raise ValueError("mass entries must be finite and positive")
    return dict(L=L, mass=mass, free=free, index=index, inverse=inverse)


def sqlap_apply(op, x, y=None, alpha=1.0, beta=0.0, mass_weight=0.0, boundary=False):
    """squared_laplacian.py:158-218 (product and Dirichlet right-hand side)."""
    L = op["L"]
    dt = np.dtype(L.dtype)
    nf = len(op["free"])
    x = np.ascontiguousarray(x, dt)
    y = np.zeros(nf, dt) if y is None else np.ascontiguousarray(y, dt)
    z = np.zeros(nf, dt)
    tmp = np.zeros(L.nrow, dt)
    lib().orc_sqlap_apply(code(dt), L.nrow, _p(L.off), _p(L.col), _p(L.val), _p(op["inverse"]),
                          _p(op["index"]), nf, _p(op["free"]), _p(op["mass"]), _p(x),
                          ctypes.c_double(0.0 if boundary else mass_weight), _p(tmp), _p(y),
                          _p(z), ctypes.c_double(alpha), ctypes.c_double(beta), int(boundary))
    return z


# ---------------------------------------------------------------- blocks
def block_jacobi(A: CSR, tolerance=None):
    """block_ilu.py:234-335 -- inverse diagonal blocks; raises on a bad pivot."""
    b = A.block
    if tolerance is None:
        tolerance = 1e-12 if A.dtype == np.float64 else 1e-6
    inv = np.zeros((A.nrow, b, b), A.dtype)
    bad = np.zeros(1, np.int32)
    vals = np.ascontiguousarray(A.val.reshape(-1, b, b))
    lib().orc_block_diagonal_only(code(A.dtype), A.nrow, b, _p(A.off), _p(A.col), None,
                                  _p(vals), _p(inv), ctypes.c_double(tolerance), _p(bad))
    if bad[0]:
        raise ValueError("Missing, nonfinite or singular/near-singular block pivot")
    return inv


def block_jacobi_apply(inv, x, y=None, alpha=1.0, beta=0.0):
    nb, b, _ = inv.shape
    x = np.ascontiguousarray(x, inv.dtype).reshape(nb, b)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, inv.dtype).reshape(nb, b)
    z = np.zeros_like(x)
    lib().orc_block_diagonal_apply(code(inv.dtype), nb, b, _p(inv), _p(x), _p(y), _p(z),
                                   ctypes.c_double(alpha), ctypes.c_double(beta))
    return z


def block_ilu0(A: CSR, factor_sweeps=12, relaxation=1.0, tolerance=None):
    """block_ilu.py:362-402 -- synchronous ILU(0) sweeps; returns (factors, inverse)."""
    b = A.block
    nb = A.nrow
    dt = code(A.dtype)
    if tolerance is None:
        tolerance = 1e-12 if A.dtype == np.float64 else 1e-6
    rows = np.zeros(A.nnz, np.int32)
    diag = np.zeros(nb, np.int32)
    bad = np.zeros(1, np.int32)
    lib().orc_block_topology(nb, _p(A.off), _p(A.col), _p(rows), _p(diag), _p(bad))
    vals = np.ascontiguousarray(A.val.reshape(-1, b, b))
    inv = np.zeros((nb, b, b), A.dtype)
    lib().orc_block_invert(dt, nb, b, _p(vals), _p(diag), _p(inv), ctypes.c_double(tolerance),
                           _p(bad))
    if bad[0]:
        raise ValueError("Missing, nonfinite or singular/near-singular block pivot")
    factors = np.zeros_like(vals)
    other = np.zeros_like(vals)
    lib().orc_ilu_initialize(dt, ctypes.c_long(A.nnz), b, _p(rows), _p(A.col), _p(vals), _p(inv),
                             _p(factors))
    for _ in range(factor_sweeps):
        lib().orc_ilu_sweep(dt, ctypes.c_long(A.nnz), b, _p(A.off), _p(A.col), _p(rows), _p(vals),
                            _p(factors), _p(inv), _p(other), ctypes.c_double(relaxation))
        factors, other = other, factors
        lib().orc_block_invert(dt, nb, b, _p(factors), _p(diag), _p(inv),
                               ctypes.c_double(tolerance), _p(bad))
    if not np.all(np.isfinite(factors)) or bad[0]:
        raise ValueError("Missing, nonfinite or singular/near-singular block pivot")
    return factors, inv


def block_ilu0_apply(A: CSR, factors, inv, x, y=None, alpha=1.0, beta=0.0, solve_sweeps=4):
    """block_ilu.py:404-450."""
    nb, b = A.nrow, A.block
    dt = code(A.dtype)
    x = np.ascontiguousarray(x, A.dtype).reshape(nb, b)
    yv = np.zeros_like(x) if y is None else np.ascontiguousarray(y, A.dtype).reshape(nb, b)
    if alpha == 0.0:
        return block_jacobi_apply(inv, x, yv, alpha, beta)
    a = np.zeros_like(x)
    bb = np.zeros_like(x)
    for _ in range(solve_sweeps):
        lib().orc_ilu_triangular(dt, nb, b, _p(A.off), _p(A.col), _p(factors), _p(inv), _p(x),
                                 _p(a), _p(bb), 1)
        a, bb = bb, a
    c = np.zeros_like(x)
    d = np.zeros_like(x)
    for _ in range(solve_sweeps):
        lib().orc_ilu_triangular(dt, nb, b, _p(A.off), _p(A.col), _p(factors), _p(inv), _p(a),
                                 _p(c), _p(d), 0)
        c, d = d, c
    out = alpha * c
    if beta != 0.0:
        out = out + beta * yv
    return out.astype(A.dtype)


def bsr_spmv(A: CSR, x, y=None, alpha=1.0, beta=0.0, vec_dtype=None):
    """Warp's bsr_mv on b x b blocks (sparse_operator.py:104): z = alpha A x + beta y,
    accumulating in the vector type, each block's product added to the row sum."""
    b = A.block
    vec_dtype = np.dtype(vec_dtype or A.dtype)
    x = np.ascontiguousarray(x, vec_dtype).reshape(-1)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, vec_dtype).reshape(-1)
    z = np.zeros(A.nrow * b, vec_dtype)
    vals = np.ascontiguousarray(A.val.reshape(-1))
    lib().orc_bsr_spmv(code(A.dtype), code(vec_dtype), A.nrow, b, _p(A.off), _p(A.col), _p(vals),
                       _p(x), _p(y), _p(z), ctypes.c_double(alpha), ctypes.c_double(beta))
    return z


# ---------------------------------------------------------------- mixed systems
def mixed_system(K: CSR, mass, order=2, data_weight=1.0, equilibrate=False):
    """mixed.py:180-243 restated: dict(matrix, H, mass, scale, data_weight).

    matrix is the interleaved order x order BSR sum of the vertex blocks
    (weight*m, -m couplings) and the K blocks (-k couplings); H = K + shift*M.
    Duplicates (K's diagonal) sum two terms, so the summation order is moot.
    """
    dt = np.dtype(K.dtype)
    n = K.nrow
    mass = np.array(mass, dt)
    scale = np.ones(order, dt)
    if equilibrate:
        root = data_weight ** (1 / order)
        scale = np.array([data_weight ** -0.25, data_weight ** 0.25] if order == 2
                         else [1 / root, 1.0, root], dt)
        mass = (dt.type(root) * mass).astype(dt)
        data_weight = 1.0
    if np.any(~np.isfinite(mass)) or np.any(mass <= 0):
        raise ValueError("mass must be finite positive")
    weight, shift = dt.type(data_weight), dt.type(data_weight ** (1 / order))
    rows = np.concatenate([np.arange(n), np.repeat(np.arange(n), np.diff(K.off))])
    cols = np.concatenate([np.arange(n), K.col])
    blocks = np.zeros((n + K.nnz, order, order), dt)
    blocks[:n, 0, 0] = weight * mass
    k = -K.val
    if order == 2:
        blocks[:n, 1, 1] = -mass
        blocks[n:, 0, 1] = blocks[n:, 1, 0] = k
    else:
        blocks[:n, 1, 2] = blocks[:n, 2, 1] = -mass
        blocks[n:, 0, 2] = blocks[n:, 2, 0] = blocks[n:, 1, 1] = k
    hvals = np.concatenate([shift * mass, K.val]).astype(dt)
    key = rows.astype(np.int64) * n + cols
    perm = np.argsort(key, kind="stable")
    uniq, first = np.unique(key[perm], return_index=True)
    vals = np.add.reduceat(blocks[perm], first, axis=0).astype(dt)
    hsum = np.add.reduceat(hvals[perm], first).astype(dt)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum(np.bincount(uniq // n, minlength=n))
    col = uniq % n
    matrix = CSR(_i32(off), _i32(col), vals, n, n, order)
    H = CSR(_i32(off), _i32(col), hsum, n, n, 1)
    return dict(matrix=matrix, H=H, K=K, mass=mass, scale=scale, order=order,
                data_weight=data_weight)


def matching_schur_apply(system, inverse, x, y=None, alpha=1.0, beta=0.0):
    """mixed.py:305-322: block elimination with repeated H inverses.

    ``inverse(v)`` applies the approximate H^-1 (e.g. ``lambda v: fsai_apply(P, v)``).
    """
    K, m, order = system["K"], system["mass"], system["order"]
    dt = m.dtype
    x = np.asarray(x, dt).reshape(-1, order)
    v = spmv(K, x[:, 1] / m)
    w = np.zeros_like(m)
    if order == 3:
        w = spmv(K, x[:, 2] / m)
        w = spmv(K, w / m)
    g = x[:, 0] - v
    if order == 3:
        g = g + w
    u = inverse(g)
    for step in range(order - 1):
        u = inverse(m * u)
        if order == 3 and step == 1:
            u = -u
    v = spmv(K, u)
    a = -(v + x[:, 1 if order == 2 else 2]) / m
    lam = np.zeros_like(m)
    if order == 3:
        v = spmv(K, a)
        lam = -(v + x[:, 1]) / m
    fields = [u, a, lam][:order]
    value = np.zeros_like(x) if alpha == 0 else np.stack(fields, 1) * dt.type(alpha)
    if beta != 0:
        value = value + dt.type(beta) * np.asarray(y, dt).reshape(-1, order)
    return value


def shifted_block_apply(system, inverse, x, y=None, alpha=1.0, beta=0.0):
    """mixed.py:330-372 and 415-427: one H inverse per field."""
    m, order = system["mass"], system["order"]
    dt = m.dtype
    x = np.asarray(x, dt).reshape(-1, order)
    scale = dt.type(system["data_weight"] ** 0.25)
    half = dt.type(0.5)
    if order == 2:
        a = (x[:, 0] / scale - scale * x[:, 1]) * half
        b = (-x[:, 0] / scale - scale * x[:, 1]) * half
        value = np.stack([inverse(a) / scale, inverse(b) * scale], 1)
    else:
        value = np.stack([inverse(-x[:, 2]), inverse(-x[:, 1]), inverse(-x[:, 0])], 1)
    value = np.zeros_like(x) if alpha == 0 else value * dt.type(alpha)
    if beta != 0:
        value = value + dt.type(beta) * np.asarray(y, dt).reshape(-1, order)
    return value


# ---------------------------------------------------------------- Krylov
def krylov(method, b, *, A: CSR | None = None, K: CSR | None = None, mass=None, weight=0.0,
           P=None, diag=None, x0=None, tol=1e-8, atol=0.0, maxiter=10000, check_every=1,
           history=False):
    """Textbook PCG (method='cg') or PCR ('cr') in float64.

    Operator: CSR ``A``, or the factored ``K diag(mass)^-1 K + weight*diag(mass)``.
    Preconditioner: FSAI dict ``P`` (G, GT), Jacobi ``diag`` or none.
    Returns (x, iterations, residual_norm[, history]).
    """
    n = len(b)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(n) if x0 is None else np.array(x0, np.float64)
    f64 = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)  # noqa: E731
    if A is not None:
        op_kind, M_ = 0, A
    else:
        op_kind, M_ = 1, K
    prec_kind = 0
    G = GT = None
    if P is not None:
        prec_kind, G, GT = 1, P["G"], P["GT"]
    elif diag is not None:
        prec_kind = 2
    hist = np.zeros(maxiter + 1) if history else None
    res = ctypes.c_double(0.0)
    it = lib().orc_krylov(
        0 if method == "cg" else 1, n, op_kind, _p(M_.off), _p(M_.col), _p(f64(M_.val)),
        _p(f64(mass)), ctypes.c_double(weight), prec_kind,
        _p(G.off if G else None), _p(G.col if G else None), _p(f64(G.val) if G else None),
        _p(GT.off if GT else None), _p(GT.col if GT else None), _p(f64(GT.val) if GT else None),
        _p(f64(diag)), _p(b), _p(x), ctypes.c_double(tol), ctypes.c_double(atol), maxiter,
        check_every, ctypes.byref(res), _p(hist),
    )
    out = (x, int(it), float(res.value))
    if history:
        return out + (hist[: it + 1],)
    return out
