"""Device plumbing: dtype ids, the current HIP stream, and GPU availability.

PyTorch-ROCm provides device memory and streams; every computation on that memory is done
by libcytnx_amd.so.
"""
from __future__ import annotations

import torch

from . import fileio

# reference dtype ids (include/Type.hpp:177-179) <-> torch dtypes
Type_C128, Type_C64, Type_F64, Type_F32 = 1, 2, 3, 4
Type_I64, Type_U64, Type_I32, Type_U32, Type_I16, Type_U16, Type_Bool = 5, 6, 7, 8, 9, 10, 11

TORCH_OF = {
    Type_C128: torch.complex128,
    Type_C64: torch.complex64,
    Type_F64: torch.float64,
    Type_F32: torch.float32,
    Type_I64: torch.int64,
    Type_U64: torch.uint64,
    Type_I32: torch.int32,
    Type_U32: torch.uint32,
    Type_I16: torch.int16,
    Type_U16: torch.uint16,
    Type_Bool: torch.bool,
}
ID_OF = {v: k for k, v in TORCH_OF.items()}
NP_OF = fileio.NP_DTYPE

CPU = -1  # Device.cpu in the reference (include/Device.hpp)


def dtype_id(t: torch.dtype) -> int:
    try:
        return ID_OF[t]
    except KeyError:
        raise TypeError("unsupported dtype %s" % t)


_COMPLEX = {Type_C128, Type_C64}
_UNSIGNED = {Type_U64, Type_U32, Type_U16, Type_Bool}
_TO_REAL = {Type_C128: Type_F64, Type_C64: Type_F32}
_TO_COMPLEX = {Type_F64: Type_C128, Type_F32: Type_C64}


def type_promote(a: int, b: int) -> int:
    """Type::type_promote (include/Type.hpp:451-470).

    Mixed complex/real promotes the real counterparts and re-complexifies; otherwise the
    smaller id wins, stepping an unsigned winner down to its signed neighbour when the other
    side is signed."""
    if a == 0 or b == 0:
        return 0
    if (a in _COMPLEX) != (b in _COMPLEX):
        r = type_promote(_TO_REAL.get(a, a), _TO_REAL.get(b, b))
        return _TO_COMPLEX.get(r, Type_C128 if r not in (Type_F32,) else Type_C64)
    if a < b:
        return a - 1 if (b not in _UNSIGNED and a in _UNSIGNED) else a
    return b - 1 if (a not in _UNSIGNED and b in _UNSIGNED) else b


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("cytnx_amd: no HIP device visible; this library runs the Device.cuda "
                           "branch only")


def torch_device(dev: int) -> torch.device:
    return torch.device("cpu") if dev < 0 else torch.device("cuda", dev)


def stream_handle(dev: int = 0) -> int:
    """hipStream_t of torch's current stream on ``dev`` (an int for ctypes)."""
    return torch.cuda.current_stream(dev).cuda_stream
