"""Dense linear-algebra front-ends on the GPU (the Device.cuda branch of cytnx::linalg).

* ``Matmul``    -- linalg::Matmul (src/linalg/Matmul.cpp:10-66) -> cxa_matmul
* ``Tensordot`` -- linalg::Tensordot (src/linalg/Tensordot.cpp:165-207) -> cxa_tensordot
* ``Gemm_Batch``-- linalg::Gemm_Batch (src/linalg/Gemm_Batch.cpp:14-236) -> cxa_gemm_batch

Validation mirrors the reference and raises before any device work.
"""
from __future__ import annotations

import ctypes
from typing import List, Sequence

import torch

from . import _lib
from .device import CPU, TORCH_OF, stream_handle, torch_device, type_promote
from .tensor import Tensor

_FLOAT_IDS = (1, 2, 3, 4)


def _as_dtype(t: Tensor, dt: int) -> Tensor:
    if t.dtype() == dt:
        return t
    c = t.contiguous()
    return Tensor.from_storage(c.storage().to(TORCH_OF[dt]), c.shape())


def _require_device(*ts: Tensor) -> int:
    dev = ts[0].device()
    for t in ts:
        if t.device() != dev:
            raise ValueError("two tensors must be on the same device")
    if dev == CPU:
        raise NotImplementedError("cytnx_amd implements the Device.cuda branch; use a GPU tensor")
    return dev


def Matmul(Tl: Tensor, Tr: Tensor) -> Tensor:
    """Row-major matrix product of rank-2 tensors."""
    if Tl.rank() != 2 or Tr.rank() != 2:
        raise ValueError("Matmul can only operate on rank-2 Tensors")
    if Tl.shape()[1] != Tr.shape()[0]:
        raise ValueError("dimension mismatch in Matmul")
    dev = _require_device(Tl, Tr)
    dt = type_promote(Tl.dtype(), Tr.dtype())
    if dt not in _FLOAT_IDS:
        raise NotImplementedError("integer Matmul is outside the GPU hot path")
    A = _as_dtype(Tl.contiguous(), dt)
    B = _as_dtype(Tr.contiguous(), dt)
    M, K = A.shape()
    N = B.shape()[1]
    out = Tensor([M, N], dt, dev, init_zero=(K == 0))
    _lib.check(_lib.lib().cxa_matmul(dt, out.storage().data_ptr(), A.storage().data_ptr(),
                                     B.storage().data_ptr(), M, K, N, stream_handle(dev)))
    return out


def _check_axes(rank: int, idx: Sequence[int], what: str) -> None:
    for i in idx:
        if i < 0 or i >= rank:
            raise ValueError("Tensordot: %s axis %d out of range" % (what, i))
    if len(set(idx)) != len(idx):
        raise ValueError("Tensordot: duplicated %s axis" % what)


def Tensordot(Tl: Tensor, Tr: Tensor, idxl: Sequence[int], idxr: Sequence[int],
              cacheL: bool = False, cacheR: bool = False) -> Tensor:
    """out[free L..., free R...] = sum over paired axes (src/linalg/Tensordot.cpp:168-187
    checks; the GEMM reads permuted operands in place where they stay strided matrices)."""
    idxl = [int(i) for i in idxl]
    idxr = [int(i) for i in idxr]
    if len(idxl) != len(idxr):
        raise ValueError("[Tensordot] error, the index list of L and R must have same length")
    if not idxl:
        raise ValueError("[Tensordot] error, pass empty index list.")
    _check_axes(Tl.rank(), idxl, "left")
    _check_axes(Tr.rank(), idxr, "right")
    for a, b in zip(idxl, idxr):
        if Tl.shape()[a] != Tr.shape()[b]:
            raise ValueError("[Tensordot] the dimension of contracted index does not match")
    dev = _require_device(Tl, Tr)
    dt = type_promote(Tl.dtype(), Tr.dtype())
    if dt not in _FLOAT_IDS:
        raise NotImplementedError("integer Tensordot is outside the GPU hot path")
    A = _as_dtype(Tl.contiguous(), dt)
    B = _as_dtype(Tr.contiguous(), dt)
    fl = [i for i in range(A.rank()) if i not in idxl]
    fr = [i for i in range(B.rank()) if i not in idxr]
    out_shape = [A.shape()[i] for i in fl] + [B.shape()[i] for i in fr]
    out = Tensor(out_shape, dt, dev, init_zero=False)
    L = _lib.lib()
    sl, sr = _lib.u64_array(A.shape()), _lib.u64_array(B.shape())
    il, ir = _lib.u64_array(idxl), _lib.u64_array(idxr)
    ws = ctypes.c_int64(0)
    _lib.check(L.cxa_tensordot_workspace(dt, A.rank(), sl, B.rank(), sr, len(idxl), il, ir,
                                         ctypes.byref(ws)))
    wbuf = torch.empty(ws.value, dtype=torch.uint8, device=torch_device(dev)) if ws.value else None
    _lib.check(L.cxa_tensordot(dt, out.storage().data_ptr(), A.storage().data_ptr(), A.rank(),
                               sl, B.storage().data_ptr(), B.rank(), sr, len(idxl), il, ir,
                               wbuf.data_ptr() if wbuf is not None else None,
                               stream_handle(dev)))
    if wbuf is not None:
        torch.cuda.current_stream(dev).synchronize()  # keep the workspace alive until done
    return out


def Gemm_Batch(alpha: Sequence[complex], a: Sequence[Tensor], b: Sequence[Tensor],
               beta: Sequence[complex], c: Sequence[Tensor], group_size: Sequence[int]) -> None:
    """C_i = alpha_g A_i B_i + beta_g C_i for row-major rank-2 tensors, in place on c
    (linalg::Gemm_Batch, src/linalg/Gemm_Batch.cpp:14-236; one grouped launch).  alpha / beta
    are the per-group Scalars: complex values for complex dtypes, real ones otherwise."""
    n = len(a)
    if not (len(b) == n and len(c) == n and sum(group_size) == n):
        raise ValueError("[Gemm_Batch] inconsistent batch sizes")
    if len(alpha) != len(group_size) or len(beta) != len(group_size):
        raise ValueError("[Gemm_Batch] alpha/beta need one entry per group")
    dev = _require_device(*a, *b, *c)
    dt = c[0].dtype()
    ms, ns, ks, lda, ldb, ldc = [], [], [], [], [], []
    idx = 0
    for g, gs in enumerate(group_size):
        for _ in range(gs):
            A, B, C = a[idx], b[idx], c[idx]
            for t in (A, B, C):
                if t.dtype() != dt or not t.is_contiguous() or t.rank() != 2:
                    raise ValueError("[Gemm_Batch] operands must be contiguous rank-2 tensors "
                                     "of one dtype")
            if A.shape()[1] != B.shape()[0] or C.shape() != [A.shape()[0], B.shape()[1]]:
                raise ValueError("[Gemm_Batch] shape mismatch in batch entry %d" % idx)
            idx += 1
    # row-major -> BLAS column-major: pass B before A, n before m (Gemm_Batch.cpp:210-230)
    flat_a, flat_b, flat_c = [], [], []
    alphas, betas, gsz = [], [], []
    idx = 0
    for g, gs in enumerate(group_size):
        for _ in range(gs):
            A, B, C = a[idx], b[idx], c[idx]
            M, K = A.shape()
            N = B.shape()[1]
            ms.append(N), ns.append(M), ks.append(K)
            lda.append(N), ldb.append(K), ldc.append(N)
            flat_a.append(B.storage().data_ptr())
            flat_b.append(A.storage().data_ptr())
            flat_c.append(C.storage().data_ptr())
            alphas.append(alpha[g])
            betas.append(beta[g])
            gsz.append(1)
            idx += 1
    cnt = len(gsz)
    VP = ctypes.c_void_p * max(1, cnt)
    if dt in (1, 2):  # complex: (re, im) per group
        def pairs(v):
            return [float(x) for z in v for x in (complex(z).real, complex(z).imag)]
        alphas, betas = pairs(alphas), pairs(betas)
    else:
        for z in list(alphas) + list(betas):
            if complex(z).imag != 0:
                raise ValueError("[Gemm_Batch] complex alpha/beta need complex operands")
        alphas = [complex(z).real for z in alphas]
        betas = [complex(z).real for z in betas]
    D = ctypes.c_double * max(1, len(alphas))
    tr = b"N" * max(1, cnt)
    _lib.check(_lib.lib().cxa_gemm_batch(
        dt, tr, tr, _lib.i64_array(ms), _lib.i64_array(ns), _lib.i64_array(ks), D(*alphas),
        VP(*flat_a), _lib.i64_array(lda), VP(*flat_b), _lib.i64_array(ldb), D(*betas),
        VP(*flat_c), _lib.i64_array(ldc), cnt, _lib.i64_array(gsz), stream_handle(dev)))


_VEC_WS = {}


def vec_workspace(dev: int) -> torch.Tensor:
    """The reduction workspace of cxa_dot / cxa_nrm2 on a device (allocated once)."""
    ws = _VEC_WS.get(dev)
    if ws is None:
        nb = _lib.lib().cxa_vec_workspace_bytes()
        ws = torch.empty(nb, dtype=torch.uint8, device=torch_device(dev))
        _VEC_WS[dev] = ws
    return ws


def _flat_data(T) -> torch.Tensor:
    from .unitensor import UniTensor
    if isinstance(T, UniTensor):
        return T.storage().reshape(-1) if T._is_block else T._block.contiguous().storage()
    return T.contiguous().storage().reshape(-1)


def Vectordot(Tl: Tensor, Tr: Tensor, is_conj: bool = False) -> Tensor:
    """linalg::Vectordot (src/linalg/Vectordot.cpp, backend Vectordotfunc_oii): sum_i Tl_i Tr_i
    (conj(Tl_i) Tr_i with is_conj) of two rank-1 tensors of equal length, as a one-element
    Tensor of the promoted dtype (cxa_dot, summed in FP64)."""
    if Tl.rank() != 1 or Tr.rank() != 1:
        raise ValueError("[Vectordot] the input Tensors should be rank-1")
    if Tl.shape()[0] != Tr.shape()[0]:
        raise ValueError("[Vectordot] the two input Tensors should have the same size")
    dev = _require_device(Tl, Tr)
    dt = type_promote(Tl.dtype(), Tr.dtype())
    if dt not in _FLOAT_IDS:
        raise NotImplementedError("integer Vectordot is outside the GPU hot path")
    a = _as_dtype(Tl.contiguous(), dt).storage()
    b = _as_dtype(Tr.contiguous(), dt).storage()
    out = torch.empty(1, dtype=TORCH_OF[dt], device=torch_device(dev))
    _lib.check(_lib.lib().cxa_dot(dt, a.numel(), a.data_ptr(), b.data_ptr(), int(bool(is_conj)),
                                  out.data_ptr(), vec_workspace(dev).data_ptr(),
                                  stream_handle(dev)))
    return Tensor.from_storage(out, [1])


def _nrm2_device(flat: torch.Tensor, dt: int, dev: int) -> torch.Tensor:
    real = {1: torch.float64, 2: torch.float32, 3: torch.float64, 4: torch.float32}[dt]
    out = torch.zeros(1, dtype=real, device=torch_device(dev))
    if flat.numel():
        _lib.check(_lib.lib().cxa_nrm2(dt, flat.numel(), flat.data_ptr(), out.data_ptr(),
                                       vec_workspace(dev).data_ptr(), stream_handle(dev)))
    return out


def norm(T) -> float:
    """linalg::norm of a Tensor or UniTensor (2-norm of all elements, cxa_nrm2), as a float."""
    flat = _flat_data(T)
    dev = T.device()
    if dev == CPU:
        raise NotImplementedError("cytnx_amd implements the Device.cuda branch; use a GPU tensor")
    dt = T.dtype()
    if dt not in _FLOAT_IDS:
        raise NotImplementedError("integer norm is outside the GPU hot path")
    return float(_nrm2_device(flat.contiguous(), dt, dev).item())


def Norm(T) -> Tensor:
    """linalg::Norm (deprecated in the reference in favour of norm): the 2-norm as a one-element
    Tensor of the real dtype, computed on the device (cxa_nrm2, Normfunc_oii)."""
    import warnings
    warnings.warn("Norm() is deprecated, use norm() instead.", DeprecationWarning, stacklevel=2)
    dev = T.device()
    if dev == CPU:
        raise NotImplementedError("cytnx_amd implements the Device.cuda branch; use a GPU tensor")
    dt = T.dtype()
    if dt not in _FLOAT_IDS:
        raise NotImplementedError("integer Norm is outside the GPU hot path")
    return Tensor.from_storage(_nrm2_device(_flat_data(T).contiguous(), dt, dev), [1])


def Gesvd(T, is_U: bool = True, is_vT: bool = True):
    """linalg::Gesvd on a UniTensor (cytnx_amd.svd)."""
    from .svd import Gesvd as _gesvd
    return _gesvd(T, is_U, is_vT)


def Svd(T, is_UvT: bool = True):
    """linalg::Svd on a UniTensor (cytnx_amd.svd)."""
    from .svd import Svd as _svd
    return _svd(T, is_UvT)


def Svd_truncate(T, keepdim: int, *args, **kwargs):
    """linalg::Svd_truncate on a UniTensor, with or without min_blockdim (cytnx_amd.svd)."""
    from .svd import Svd_truncate as _svdt
    return _svdt(T, keepdim, *args, **kwargs)


def Gesvd_truncate(T, keepdim: int, *args, **kwargs):
    """linalg::Gesvd_truncate on a UniTensor, with or without min_blockdim (cytnx_amd.svd)."""
    from .svd import Gesvd_truncate as _gsvdt
    return _gsvdt(T, keepdim, *args, **kwargs)


def Lanczos(*args, **kwargs):
    """linalg::Lanczos: method="Gnd", or which="SA"/"LA"/"LM" (cytnx_amd.lanczos.Lanczos)."""
    from .lanczos import Lanczos as _lanczos
    return _lanczos(*args, **kwargs)


def Lanczos_Exp(*args, **kwargs):
    """linalg::Lanczos_Exp for a real or complex tau (cytnx_amd.lanczos.Lanczos_Exp)."""
    from .lanczos import Lanczos_Exp as _lanczos_exp
    return _lanczos_exp(*args, **kwargs)


__all__ = ["Matmul", "Tensordot", "Gemm_Batch", "Vectordot", "norm", "Norm", "Gesvd", "Svd", "Svd_truncate",
           "Gesvd_truncate", "Lanczos", "Lanczos_Exp"]
