"""Matrix Market dumps in the layout the reference's dragon benchmark reads:
host-side parsing and consistency checks (device upload is covered on the GPU)."""

import numpy as np
import pytest
import scipy.io as sio
import scipy.sparse as ss

from warp_preconditioners import ingest, workloads


def _write_dump(root, corrupt=False):
    v, t, _ = workloads.kuhn_grid(5, 0.15, seed=1)
    K, m = workloads.p1_laplacian(v, t)
    n = len(m)
    Q = workloads.squared_operator(K, m, 1.0)
    H = (ss.diags(m) + K).tocsr()
    mixed = ss.bmat([[ss.diags(m), -K], [-K, -ss.diags(m)]]).tocsr()
    sio.mmwrite(root / "k1_Q.mtx", H if not corrupt else 2 * H)
    sio.mmwrite(root / "k2_Q.mtx", Q)
    sio.mmwrite(root / "k4_Q.mtx", mixed)
    sio.mmwrite(root / "k2_rhs.mtx", np.random.default_rng(0).standard_normal((n, 3)))
    return K, m, Q


def test_dump_checks_pass_and_match(tmp_path, monkeypatch):
    K, m, Q = _write_dump(tmp_path)
    monkeypatch.setattr(ingest, "upload", lambda a, device=None, dtype=None: a)
    monkeypatch.setattr(ingest.torch, "tensor", lambda a, **kw: np.asarray(a))
    d = ingest.load_dump(tmp_path, 2, device="cpu")
    assert d["rhs"].shape == (len(m), 3)
    np.testing.assert_allclose(d["mass_host"], m)
    assert abs(d["K_host"] - K).max() == 0
    assert abs(d["Q"] - Q).max() < 1e-12 * abs(Q).max()


def test_dump_inconsistency_raises(tmp_path):
    _write_dump(tmp_path, corrupt=True)
    with pytest.raises(ValueError, match="M \\+ K"):
        ingest.load_dump(tmp_path, 2, device="cpu")


def test_backward_error_definition():
    a = ss.csr_matrix(np.array([[2.0, -1.0], [-1.0, 2.0]]))
    x = np.array([1.0, 1.0])
    b = a @ x
    assert ingest.backward_error(a, abs(a), b, x) == 0.0
    b2 = b + np.array([3e-3, 0.0])
    assert np.isclose(ingest.backward_error(a, abs(a), b2, x), 3e-3 / (3.0 + 1.003))


def _beam(fixed_x0=True, seed=2):
    v, t, _ = workloads.kuhn_grid(None, 0.15, seed=seed, shape=(8, 3, 3))
    fixed = (v[:, 0] == 0.0) if fixed_x0 else np.zeros(len(v), dtype=bool)
    return v, t, fixed


def test_p1_elasticity_closed_forms():
    """With no vertex fixed the stiffness is symmetric positive semidefinite and
    annihilates the six rigid motions (three translations, three infinitesimal
    rotations u = w x X); the lumped volumes sum to the mesh volume."""
    v, t, fixed = _beam(fixed_x0=False)
    K, rhs, mass, free = workloads.p1_elasticity(v, t, fixed, young=3.0, poisson=0.3)
    A = K.toarray()
    assert abs(A - A.T).max() <= 1e-13 * abs(A).max()
    assert np.linalg.eigvalsh(A).min() >= -1e-11 * abs(A).max()
    for c in range(3):
        u = np.zeros_like(v)
        u[:, c] = 1.0
        assert abs(A @ u.reshape(-1)).max() <= 1e-12 * abs(A).max()
        w = np.zeros(3)
        w[c] = 1.0
        rot = np.cross(w, v)
        assert abs(A @ rot.reshape(-1)).max() <= 1e-11 * abs(A).max()
    # a pure shear strain u = (y, 0, 0) is not rigid: it has positive energy mu*|V|
    shear = np.zeros_like(v)
    shear[:, 0] = v[:, 1]
    mu = 3.0 / (2 * 1.3)
    assert np.isclose(shear.reshape(-1) @ A @ shear.reshape(-1), mu * mass.sum(), rtol=1e-10)
    # a uniform dilatation u = X: strain I, energy density 2 mu tr(I) + lam tr(I)^2
    lam = 3.0 * 0.3 / (1.3 * 0.4)
    dil = v.reshape(-1)
    assert np.isclose(dil @ A @ dil, (6 * mu + 9 * lam) * mass.sum(), rtol=1e-10)
    lx, ly, lz = np.ptp(v, axis=0)
    assert np.isclose(mass.sum(), lx * ly * lz, rtol=1e-12)


def test_mesh_dump_material_and_loads(tmp_path, monkeypatch):
    import json

    monkeypatch.setattr(ingest, "from_scipy", lambda a, device=None, dtype=None, bs=None: a)
    monkeypatch.setattr(ingest.torch, "tensor", lambda a, **kw: np.asarray(a))
    v, t, fixed = _beam()
    # defaults of the reference benchmark, gravity load
    np.savez(tmp_path / "mesh.npz", vertices=v, tets=t.astype(np.int32), fixed=fixed)
    d = ingest.load_mesh_npz(tmp_path, device="cpu")
    K0, _, mass, free = workloads.p1_elasticity(v, t, fixed, young=1e7, poisson=0.35)
    assert abs(d["K_host"] - K0).max() == 0.0
    np.testing.assert_array_equal(d["free"], np.flatnonzero(~fixed))
    np.testing.assert_allclose(d["rhs"], 1000.0 * mass[free][:, None] * [0.0, 0.0, -9.81],
                               rtol=1e-15)
    assert d["material"] == {"young_pa": 1e7, "poisson": 0.35, "density_kg_m3": 1000.0}
    # material from mesh.json, stored nodal forces win over gravity
    forces = np.random.default_rng(4).standard_normal(v.shape)
    np.savez(tmp_path / "mesh.npz", vertices=v, tets=t, fixed=fixed, forces=forces)
    (tmp_path / "mesh.json").write_text(json.dumps({"young_pa": 2.0, "poisson": 0.45}))
    d = ingest.load_mesh_npz(tmp_path, device="cpu")
    K1, _, _, _ = workloads.p1_elasticity(v, t, fixed, young=2.0, poisson=0.45)
    assert abs(d["K_host"] - K1).max() == 0.0
    np.testing.assert_array_equal(d["rhs"], forces[~fixed])
    np.testing.assert_array_equal(d["rhs_device"], forces[~fixed].reshape(-1))
    np.savez(tmp_path / "mesh.npz", vertices=v, tets=t[:, :3], fixed=fixed)
    with pytest.raises(ValueError, match="tets"):
        ingest.load_mesh_npz(tmp_path, device="cpu")
