"""Device assembly against the host generators: P1 stiffness and mass, and the
squared operator, which must be bit-identical to the host SpGEMM for the same K."""

import numpy as np
import pytest
import scipy.sparse as ss
import torch

from gpu_helpers import dev, host, upload

pytestmark = pytest.mark.gpu


def test_p1_assembly_matches_host():
    from warp_preconditioners import assembly, workloads

    v, t, _ = workloads.kuhn_grid(12, 0.15, seed=3)
    K, m = workloads.p1_laplacian(v, t)
    tets = assembly.kuhn_tets((12, 12, 12), dev())
    np.testing.assert_array_equal(host(tets), t)
    Kd, md = assembly.p1_laplacian(torch.tensor(v, device=dev()), tets)
    np.testing.assert_array_equal(host(Kd.offsets), K.indptr)
    np.testing.assert_array_equal(host(Kd.columns), K.indices)
    np.testing.assert_allclose(host(Kd.values), K.data, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(host(md), m, rtol=1e-13)
    # determinism: a second assembly is bit-identical
    Kd2, md2 = assembly.p1_laplacian(torch.tensor(v, device=dev()), tets)
    np.testing.assert_array_equal(host(Kd2.values), host(Kd.values))
    np.testing.assert_array_equal(host(md2), host(md))


@pytest.mark.parametrize("alpha", [1.0, 1e-4])
def test_squared_operator_bitwise_equals_host_spgemm(alpha):
    from warp_preconditioners import assembly, workloads

    v, t, _ = workloads.kuhn_grid(16, 0.15, seed=0)
    K, m = workloads.p1_laplacian(v, t)
    Q = workloads.squared_operator(K, m, alpha)
    Qd = assembly.squared_operator(upload(K), torch.tensor(m, device=dev()), alpha)
    # The device product is structural: it keeps every column reachable through a
    # stored entry of K (this grid's K stores exact zeros on boundary faces), where
    # scipy drops products that are exactly zero.
    P = K.copy()
    P.data[:] = 1.0
    S = (P @ P).tocsr()
    S.sort_indices()
    np.testing.assert_array_equal(host(Qd.offsets), S.indptr)
    np.testing.assert_array_equal(host(Qd.columns), S.indices)
    D = ss.csr_matrix((host(Qd.values), host(Qd.columns), host(Qd.offsets)), shape=Q.shape)
    D.eliminate_zeros()
    np.testing.assert_array_equal(D.indptr, Q.indptr)
    np.testing.assert_array_equal(D.indices, Q.indices)
    np.testing.assert_array_equal(D.data, Q.data)


def test_device_pipeline_factor_matches_oracle_hash():
    # The full 64^3 biharmonic Q assembled from the host K on the device: the FSAI
    # factor equals the oracle's (golden SHA-256).
    import hashlib
    import json
    from pathlib import Path

    from warp_preconditioners import FSAI, assembly, workloads

    meta = json.loads((Path(__file__).parent / "golden" / "patterns.json").read_text())
    v, t, _ = workloads.kuhn_grid(64, 0.15, seed=0)
    K, m = workloads.p1_laplacian(v, t)
    Qd = assembly.squared_operator(upload(K), torch.tensor(m, device=dev()), 1.0)
    P = FSAI(Qd, max_row_size=8, factor_dtype=torch.float32)
    h = hashlib.sha256()
    h.update(host(P.G.offsets).tobytes())
    h.update(host(P.G.columns).tobytes())
    assert h.hexdigest() == meta["biharmonic64_Q_w8"]["pattern_sha256"]
    assert hashlib.sha256(host(P.G.values).tobytes()).hexdigest() == \
        meta["biharmonic64_Q_w8"]["values_sha256"]


@pytest.mark.parametrize("side", [16, 128])
def test_device_spgemm_reproduces_recorded_grid_hashes(side):
    # The device SpGEMM (K diag(1/m) K + alpha M with K = the five-point Laplacian,
    # m = 1, alpha = 0.05) builds the reference's grid() matrix byte for byte: the
    # four matrix_sha256 values of results/fsai-float16-regression.json (float32 by
    # the device conversion of the float64 values).
    from test_golden import GRID_SHA256, grid_signature
    from warp_preconditioners import assembly, workloads
    from warp_preconditioners.sparse import bsr_copy

    L = upload(workloads.laplacian_2d(side))
    ones = torch.ones(side * side, dtype=torch.float64, device=dev())
    Q = assembly.squared_operator(L, ones, 0.05)
    for dt, name in ((torch.float64, "float64"), (torch.float32, "float32")):
        Qd = Q if dt == torch.float64 else bsr_copy(Q, scalar_type=dt)
        got = grid_signature(host(Qd.offsets), host(Qd.columns), host(Qd.values))
        assert got == GRID_SHA256[(side, name)]
