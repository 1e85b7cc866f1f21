"""Benchmark configurations on the GPU: bit-exact factors against the oracle's golden
fixtures (full config 1; SHA-256 of the full-size factors of configs 2-4 plus 1000
sampled rows), and iteration parity of the solvers."""

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from gpu_helpers import dev, host, upload, vec

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_laplace2d64_factor_bitwise_and_cg(oracle):
    from warp_preconditioners import FSAI, linear, workloads

    g = np.load(GOLDEN / "laplace2d64_fsai_w8.npz")
    lap = workloads.laplacian_2d(64)
    A = upload(lap, torch.float32)
    P = FSAI(A, max_row_size=8)
    np.testing.assert_array_equal(host(P.G.offsets), g["offsets"])
    np.testing.assert_array_equal(host(P.G.columns), g["columns"])
    np.testing.assert_array_equal(host(P.G.values), g["values"])
    b = np.random.default_rng(0).standard_normal(4096)
    # float64 vectors with the float32 factor: the oracle's exact count
    A64 = upload(lap)
    x = torch.zeros(4096, dtype=torch.float64, device=dev())
    it, res, atol = linear.cg(A64, vec(b), x, M=P, tol=1e-8, maxiter=4096, check_every=1)
    assert it == int(g["cg_iterations"])
    # float32 plain path (the configuration as stated): within 2% of the count
    x32 = torch.zeros(4096, dtype=torch.float32, device=dev())
    it32, _, _ = linear.cg(A, vec(b, torch.float32), x32, M=P, tol=1e-8, maxiter=4096,
                           check_every=1)
    assert abs(it32 - int(g["cg_iterations"])) <= max(1, 0.02 * int(g["cg_iterations"]))


def _workload(name):
    from warp_preconditioners import workloads
    from warp_preconditioners.sparse import BsrMatrix

    if name == "biharmonic64_Q_w8":
        v, t, _ = workloads.kuhn_grid(64, 0.15, seed=0)
        K, m = workloads.p1_laplacian(v, t)
        return upload(workloads.squared_operator(K, m, 1.0))
    if name == "kron48_KxC_w16":
        off, col, val, _, _ = workloads.kron_coupling(48, 1.0e6, seed=0)
        return BsrMatrix(len(off) - 1, len(off) - 1, (3, 3), torch.tensor(off, device=dev()),
                         torch.tensor(col, device=dev()), torch.tensor(val, device=dev()))
    vb, tb, _ = workloads.kuhn_grid(None, 0.15, seed=0, shape=(160, 40, 40))
    Kb, _, _, _ = workloads.p1_elasticity(vb, tb, vb[:, 0] == 0.0, young=1.0, poisson=0.45)
    return BsrMatrix(Kb.shape[0] // 3, Kb.shape[1] // 3, (3, 3),
                     torch.tensor(Kb.indptr.astype(np.int32), device=dev()),
                     torch.tensor(Kb.indices.astype(np.int32), device=dev()),
                     torch.tensor(np.ascontiguousarray(Kb.data), device=dev()))


@pytest.mark.parametrize("name", ["biharmonic64_Q_w8", "kron48_KxC_w16", "beam40x40x160_elasticity_w8"])
def test_full_size_factor_matches_oracle_hash(name):
    from warp_preconditioners import FSAI

    meta = json.loads((GOLDEN / "patterns.json").read_text())[name]
    A = _workload(name)
    fd = {"float32": torch.float32, "float64": torch.float64}[meta["fsai"]["factor_dtype"]]
    P = FSAI(A, max_row_size=meta["fsai"]["max_row_size"], factor_dtype=fd)
    off, col, val = host(P.G.offsets), host(P.G.columns), host(P.G.values)
    assert P.G.nnz == meta["nnz"] and P.truncated_rows == meta["truncated_rows"]
    assert digest(off, col) == meta["pattern_sha256"]
    rows = np.load(GOLDEN / f"rows_{name}.npz")
    for k, r in enumerate(rows["rows"]):
        a, b = off[r], off[r + 1]
        np.testing.assert_array_equal(col[a:b], rows["cols"][rows["off"][k]:rows["off"][k + 1]])
        np.testing.assert_array_equal(val[a:b], rows["vals"][rows["off"][k]:rows["off"][k + 1]])
    assert digest(val) == meta["values_sha256"]
