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
