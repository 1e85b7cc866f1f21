"""The oracle reproduces the committed golden fixtures (regression pin)."""

import json
from pathlib import Path

import numpy as np

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
