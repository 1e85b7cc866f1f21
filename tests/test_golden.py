"""The oracle reproduces the committed golden fixtures (regression pin)."""

import json
from pathlib import Path

import numpy as np
import scipy.sparse as ss

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_laplace2d64_factor_and_iterations(oracle):
    from warp_preconditioners import workloads

    g = np.load(GOLDEN / "laplace2d64_fsai_w8.npz")
    lap = workloads.laplacian_2d(64)
    P = oracle.fsai(oracle.CSR.from_scipy(lap, np.float32), max_row_size=8)
    np.testing.assert_array_equal(P["G"].off, g["offsets"])
    np.testing.assert_array_equal(P["G"].col, g["columns"])
    np.testing.assert_array_equal(P["G"].val, g["values"])
    assert P["truncated_rows"] == int(g["truncated"])
    b = np.random.default_rng(0).standard_normal(4096)
    P64 = dict(G=oracle.CSR(P["G"].off, P["G"].col, P["G"].val.astype(np.float64), 4096, 4096),
               GT=oracle.CSR(P["GT"].off, P["GT"].col, P["GT"].val.astype(np.float64), 4096, 4096))
    x, it, _ = oracle.krylov("cg", b, A=oracle.CSR.from_scipy(lap), P=P64, tol=1e-8, maxiter=4096)
    assert it == int(g["cg_iterations"])
    assert np.linalg.norm(b - lap @ x) <= 1e-8 * np.linalg.norm(b)


def test_pattern_fixtures_are_consistent():
    meta = json.loads((GOLDEN / "patterns.json").read_text())
    for name, m in meta.items():
        rows = np.load(GOLDEN / f"rows_{name}.npz")
        assert len(rows["rows"]) == 1000
        assert rows["off"][-1] == len(rows["cols"]) == len(rows["vals"])
        for k, r in enumerate(rows["rows"][:50]):
            cols = rows["cols"][rows["off"][k]:rows["off"][k + 1]]
            assert cols[-1] == r and np.all(np.diff(cols) > 0)  # diagonal last, sorted
        assert m["nnz"] <= m["n"] * m["fsai"]["max_row_size"]


# matrix_sha256 of the reference's grid() matrices L L + 0.05 I (L: the 2D five-point
# Laplacian with diagonal 4 on a side x side grid), offsets int32 + columns int32 +
# values, as recorded in results/fsai-float16-regression.json (generator:
# benchmarks/fsai_dtype_regression.py:57-75, hash: :48-54).
GRID_SHA256 = {
    (16, "float32"): "25738865eb35d5547b0f02d5ff69581c239330a6d85c52526d23bf81021b135b",
    (16, "float64"): "3d9c9091a41b78563009737d63940856cf1b6e1a9332dc7e79caf8e415ed45ad",
    (128, "float32"): "680991d20bff6d8a9f91933d5e8c223a5336a17997d7d4566424ed0df77e9633",
    (128, "float64"): "c4764eb87e4b70bc535c928b200f31930f4c2089a3e1e839a5cf99b3a858128f",
}


def grid_signature(offsets, columns, values):
    import hashlib

    digest = hashlib.sha256()
    for a in (offsets.astype(np.int32), columns.astype(np.int32), values):
        digest.update(np.ascontiguousarray(a).tobytes())
    return digest.hexdigest()


def test_laplacian_2d_generator_reproduces_recorded_grid_hashes():
    # The configuration-1 generator (workloads.laplacian_2d), squared and shifted
    # by 0.05 I exactly as the reference's grid() does, hashes to the four recorded
    # values (a byte-level pin of the generator's pattern, ordering and values).
    from warp_preconditioners import workloads

    for (side, dt), want in GRID_SHA256.items():
        lap = workloads.laplacian_2d(side)
        q = (lap @ lap).tocsr()
        q.sort_indices()
        q = (q + 0.05 * ss.eye(side * side)).tocsr()
        q.sort_indices()
        assert q.nnz == (3012 if side == 16 else 210436)
        assert grid_signature(q.indptr, q.indices, q.data.astype(dt)) == want
