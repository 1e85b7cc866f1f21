"""Matrix Market ingest onto the device (benchmarks/dragon.py:79-96 layout): the
uploaded matrices equal the dump's, FSAI-CG on the uploaded Q and on the factored
operator built from the uploaded K and mass both solve the dump's right-hand
sides, and the componentwise backward error (benchmarks/dragon.py:34-37) of each
result against the host Q is at the level the solver's tolerance allows."""

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host

pytestmark = pytest.mark.gpu


def _write_dump(root, side=7):
    from warp_preconditioners import workloads

    v, t, _ = workloads.kuhn_grid(side, 0.15, seed=3)
    K, m = workloads.p1_laplacian(v, t)
    n = len(m)
    Q = workloads.squared_operator(K, m, 1.0)
    H = (ss.diags(m) + K).tocsr()
    mixed = ss.bmat([[ss.diags(m), -K], [-K, -ss.diags(m)]]).tocsr()
    sio.mmwrite(root / "k1_Q.mtx", H)
    sio.mmwrite(root / "k2_Q.mtx", Q)
    sio.mmwrite(root / "k4_Q.mtx", mixed)
    rhs = np.random.default_rng(11).standard_normal((n, 3))
    sio.mmwrite(root / "k2_rhs.mtx", rhs)
    return K, m, Q, rhs


def test_dump_uploads_and_solves(tmp_path):
    import warp_preconditioners as wp
    from warp_preconditioners import ingest, linear

    K, m, Q, rhs = _write_dump(tmp_path)
    d = ingest.load_dump(tmp_path, 2)
    n = Q.shape[0]
    assert d["A"].device.type == "cuda" and d["K"].device.type == "cuda"
    got_q = ss.csr_matrix((host(d["A"].values).reshape(-1), host(d["A"].columns),
                           host(d["A"].offsets)), shape=(n, n))
    assert abs(got_q - Q).max() <= 1e-12 * abs(Q).max()
    got_k = ss.csr_matrix((host(d["K"].values).reshape(-1), host(d["K"].columns),
                           host(d["K"].offsets)), shape=(n, n))
    assert abs(got_k - K).max() == 0.0
    np.testing.assert_allclose(host(d["mass"]), m, rtol=1e-15)
    np.testing.assert_array_equal(host(d["rhs_device"]), rhs)

    P = wp.FSAI(d["A"], max_row_size=8)
    qa = abs(Q)
    ops = {
        "assembled": linear.aslinearoperator(d["A"]),
        "factored": wp.SquaredLaplacianOperator(
            d["K"], d["mass"], torch.arange(n, dtype=torch.int32, device=dev()),
            mass_weight=1.0),
    }
    for name, op in ops.items():
        for j in range(rhs.shape[1]):
            b = d["rhs_device"][:, j].contiguous()
            x = torch.zeros(n, dtype=torch.float64, device=dev())
            it, _, _ = linear.cg(op, b, x, tol=1e-12, maxiter=5000, M=P, check_every=10)
            xh = host(x)
            rel = np.linalg.norm(rhs[:, j] - Q @ xh) / np.linalg.norm(rhs[:, j])
            assert rel <= 1e-10, (name, j, it, rel)
            berr = ingest.backward_error(Q, qa, rhs[:, j], xh)
            assert berr <= 1e-12, (name, j, berr)


def test_mesh_dump_uploads_and_solves(tmp_path):
    """A mesh dump in the layout of the reference's elasticity benchmarks (mesh.npz +
    mesh.json) assembled on the free vertices as 3x3 BSR on the device; FSAI-CG and
    block-Jacobi-CG on it reach the requested residual against the host stiffness."""
    import json

    import warp_preconditioners as wp
    from warp_preconditioners import ingest, linear, workloads

    v, t, _ = workloads.kuhn_grid(None, 0.15, seed=5, shape=(16, 5, 5))
    fixed = v[:, 0] == 0.0
    np.savez(tmp_path / "mesh.npz", vertices=v, tets=t, fixed=fixed)
    (tmp_path / "mesh.json").write_text(json.dumps({"young_pa": 1.0, "poisson": 0.45,
                                                   "density_kg_m3": 1.0}))
    d = ingest.load_mesh_npz(tmp_path)
    A, K = d["A"], d["K_host"]
    assert A.block_shape == (3, 3) and A.nrow == len(d["free"])
    dense = A.to_dense()
    assert abs(dense - K.toarray()).max() == 0.0
    b = d["rhs_device"]
    bh = d["rhs"].reshape(-1)
    for name, P in (("fsai", wp.FSAI(A, max_row_size=8)), ("jacobi", wp.BlockJacobi(A))):
        x = torch.zeros_like(b)
        it, _, _ = linear.cg(linear.aslinearoperator(A), b, x, tol=1e-10, maxiter=20000, M=P,
                             check_every=10)
        rel = np.linalg.norm(bh - K @ host(x)) / np.linalg.norm(bh)
        assert rel <= 2e-10, (name, it, rel)
