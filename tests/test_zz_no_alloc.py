"""The allocation-free FSAI apply (wpc_fsai_apply_noalloc) against wpc_fsai_apply
with a caller temporary, through the C ABI: bit-identical on two calls in a row,
the second with a smaller input, and the size limit refused before any launch."""

import numpy as np
import pytest
import torch

from gpu_helpers import dev, host, random_spd, upload

pytestmark = pytest.mark.gpu

WPC_ESCRATCH = -3
LIMIT = 32 << 20  # WPC_NOALLOC_SCRATCH_BYTES


def _args(P, x, y, z, nvec, alpha, beta, lanes):
    from warp_preconditioners import _native as nat

    G, GT = P.G, P.GT
    return [G.nrow, nat.ptr(G.offsets), nat.ptr(G.columns), nat.ptr(G.values),
            nat.ptr(GT.offsets), nat.ptr(GT.columns), nat.ptr(GT.values),
            nat.code(G.scalar_type), nat.ptr(P.row_scale2), nat.ptr(x), nat.ptr(y), nat.ptr(z),
            nat.code(x.dtype), nvec, float(alpha), float(beta), lanes, nat.stream()]


def _both(P, nvec, dtype, alpha, beta, lanes, seed):
    """(z from wpc_fsai_apply with a caller temporary, z from the allocation-free
    entry point) on the same seeded x and y."""
    from warp_preconditioners import _native as nat

    n = P.G.nrow
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.normal(size=n * nvec), dtype=dtype, device=dev())
    y = torch.tensor(rng.normal(size=n * nvec), dtype=dtype, device=dev())
    z_ref = torch.full_like(x, np.nan)
    z_new = torch.full_like(x, np.nan)
    tmp = torch.empty(n * nvec, dtype=dtype, device=dev())
    a = _args(P, x, y, z_ref, nvec, alpha, beta, lanes)
    assert nat.lib().wpc_fsai_apply(*a[:9], nat.ptr(tmp), *a[9:]) == 0
    assert nat.lib().wpc_fsai_apply_noalloc(*_args(P, x, y, z_new, nvec, alpha, beta, lanes)) == 0
    torch.cuda.synchronize()
    return host(z_ref), host(z_new)


def test_noalloc_apply_matches_caller_temporary_twice():
    import warp_preconditioners as wp

    rng = np.random.default_rng(5)
    P1 = wp.FSAI(upload(random_spd(rng, 4000, 0.003)), max_row_size=8)
    # first call: 4,000 rows x 3 float64 vectors
    ref, got = _both(P1, 3, torch.float64, 0.7, -0.3, 8, 1)
    assert np.isfinite(got).all()
    np.testing.assert_array_equal(got, ref)
    # second, smaller: 900 rows x 1 float32 vector, row-scaled bfloat16 factor (the
    # array still holds the first call's larger contents)
    small = upload(random_spd(np.random.default_rng(6), 900, 0.01), torch.float32)
    P2 = wp.FSAI(small, max_row_size=16, factor_dtype=torch.bfloat16)
    ref, got = _both(P2, 1, torch.float32, 1.0, 0.0, 4, 2)
    assert np.isfinite(got).all()
    np.testing.assert_array_equal(got, ref)
    # alpha = 0: neither x nor the temporary is read
    ref, got = _both(P1, 3, torch.float64, 0.0, 2.0, 8, 3)
    np.testing.assert_array_equal(got, ref)


def test_noalloc_apply_refuses_oversized_temporary():
    from warp_preconditioners import _native as nat

    # one float64 row past the limit, 1 and 8 vectors: refused before any pointer
    # is read or anything is launched
    for nvec in (1, 8):
        n = LIMIT // (8 * nvec) + 1
        rc = nat.lib().wpc_fsai_apply_noalloc(n, None, None, None, None, None, None, 3, None,
                                              None, None, None, 4, nvec, 1.0, 0.0, 8, None)
        assert rc == WPC_ESCRATCH
