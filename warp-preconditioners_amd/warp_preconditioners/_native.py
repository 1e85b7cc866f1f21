"""ctypes binding of the gfx950 HIP library (libwpc.so, declared in include/wpc.h).

The library is loaded after ``import torch`` so that it binds to the same HIP
runtime as torch's device buffers.  There is no fallback: if the library is
missing or no HIP device is present, every product entry point raises.
"""

from __future__ import annotations

import ctypes
import re
import threading
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "_native" / "libwpc.so"

F16, BF16, F32, F64 = 1, 2, 3, 4
_DTYPE_CODES = {torch.float16: F16, torch.bfloat16: BF16, torch.float32: F32, torch.float64: F64}

HEADER = Path(__file__).resolve().parents[2] / "include" / "wpc.h"


def _parse_header(path: Path) -> dict:
    """Read the C ABI from include/wpc.h: name -> argument codes (i, l, d, p)."""
    text = re.sub(r"/\*.*?\*/", " ", path.read_text(), flags=re.S)
    out = {}
    for ret, name, args in re.findall(r"\b(int|const char\*)\s+(wpc_\w+)\s*\(([^)]*)\)\s*;", text):
        codes = ""
        for arg in args.split(","):
            arg = arg.strip()
            if not arg or arg == "void":
                continue
            if "*" in arg:
                codes += "p"
            elif arg.startswith("double"):
                codes += "d"
            elif arg.startswith("long"):
                codes += "l"
            elif arg.startswith("int"):
                codes += "i"
            else:
                raise ValueError(f"unsupported ABI argument {arg!r} in {name}")
        out[name] = (codes, ret)
    return out


_SIGNATURES = _parse_header(HEADER)
_CT = {"i": ctypes.c_int, "l": ctypes.c_long, "d": ctypes.c_double, "p": ctypes.c_void_p}

_lib = None


class NativeError(RuntimeError):
    pass


def available() -> bool:
    return LIB_PATH.exists()


def lib():
    """Load libwpc.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise NativeError(
                f"{LIB_PATH} is missing; build it with `python warp-preconditioners_amd/build.py`"
            )
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (sig, ret) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = [_CT[c] for c in sig]
            fn.restype = ctypes.c_char_p if ret != "int" else ctypes.c_int
        _lib = handle
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def code(dtype) -> int:
    try:
        return _DTYPE_CODES[dtype]
    except KeyError:
        raise TypeError(f"unsupported dtype {dtype}") from None


def ptr(t):
    return None if t is None else t.data_ptr()


# Counts reallocations of operator-owned scratch buffers (product partials, the
# two-pass temporaries).  A captured graph bakes in the addresses of such buffers,
# so a graph recorded at one generation is replayed only while the generation is
# unchanged; after a reallocation the old buffer may belong to another tensor.
_generation = 0


def bump_generation() -> None:
    global _generation
    _generation += 1


def generation() -> int:
    return _generation


def stream():
    return torch.cuda.current_stream().cuda_stream


class PrivateStream:
    """A HIP stream of one owner's own (wpc_stream_create), usable as a torch stream
    (``.torch``).  torch hands out streams from a pool of 32 per device, so two
    torch streams can be the same HIP stream; a stream that graphs are captured on
    must be one that no other work (another thread's, another system's) runs on."""

    def __init__(self, device=None):
        handle = ctypes.c_void_p()
        call("wpc_stream_create", ctypes.addressof(handle))
        self.handle = handle.value
        self.torch = torch.cuda.ExternalStream(self.handle, device=device)

    def release(self):
        if self.handle is not None:
            lib().wpc_stream_destroy(self.handle)
            self.handle = None


_capture_streams = threading.local()


def capture_stream() -> torch.cuda.Stream:
    """This thread's private capture stream on the current device (created once,
    kept for the process)."""
    dev = torch.cuda.current_device()
    cache = getattr(_capture_streams, "by_device", None)
    if cache is None:
        cache = _capture_streams.by_device = {}
    if dev not in cache:
        cache[dev] = PrivateStream(dev)
    return cache[dev].torch


def require_device(t: torch.Tensor, what: str = "array"):
    if not t.is_cuda:
        raise ValueError(f"{what} must live on a HIP device (got {t.device})")


def call(name: str, *args) -> None:
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        meaning = {-1: "unsupported dtype combination", -2: "argument out of range",
                   -3: "workspace too small"}.get(rc, f"HIP error {rc}")
        raise NativeError(f"{name} failed: {meaning}")
