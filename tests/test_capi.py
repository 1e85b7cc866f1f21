"""The C-ABI library loads and exports every symbol include/wpc.h declares (no GPU
calls), and the ctypes layer refuses to run without a HIP device."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared():
    text = re.sub(r"/\*.*?\*/", " ", (ROOT / "include" / "wpc.h").read_text(), flags=re.S)
    return re.findall(r"\b(?:int|const char\*)\s+(wpc_\w+)\s*\(", text)


def test_library_exports_every_declared_symbol():
    from warp_preconditioners import _native

    if not _native.available():
        pytest.skip("libwpc.so not built")
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(names) == sorted(_native.exported_symbols())


def test_version_string_without_device():
    from warp_preconditioners import _native

    if not _native.available():
        pytest.skip("libwpc.so not built")
    assert b"gfx950" in _native.lib().wpc_version()


def test_product_refuses_host_tensors():
    import numpy as np
    import torch

    import warp_preconditioners as wp
    from warp_preconditioners.sparse import BsrMatrix

    a = BsrMatrix(2, 2, (1, 1), torch.tensor([0, 1, 2], dtype=torch.int32),
                  torch.tensor([0, 1], dtype=torch.int32), torch.tensor([1.0, 2.0]))
    with pytest.raises(ValueError, match="HIP device"):
        wp.FSAI(a)
    # argument validation happens before any device work
    with pytest.raises(ValueError, match="max_row_size"):
        wp.FSAI(a, max_row_size=0)
    with pytest.raises(ValueError, match="apply_lanes"):
        wp.FSAI(a, apply_lanes=3)
    with pytest.raises(ValueError, match="factor_dtype"):
        wp.FSAI(a, factor_dtype=torch.float64)
    assert np.isfinite(1.0)
