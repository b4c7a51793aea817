"""LinOp and ground-state Lanczos (the reference's linalg::Lanczos(method="Gnd")).

``Lanczos(Hop, Tin, method="Gnd", CvgCrit, Maxiter)`` mirrors the Python binding of
src/linalg/Lanczos.cpp:135-179 / Lanczos_Gnd_Ut.cpp: it returns ``[energy, eigvector]``.
The iteration itself runs in the C++ driver (cxa_lanczos_run) on device vectors:

* for the planned effective Hamiltonians (cytnx_amd.heff) the matvec is the C-side chain, so an
  iteration has no Python in it;
* any other LinOp subclass is called back through Python per matvec, like the reference's
  PyLinOp trampoline (pybind/linop_py.cpp:22-41).

``ground_state_mixed`` is the FP32 + refinement scheme: Lanczos with the contraction bulk on
FP32 MFMA (vectors, scalars and re-orthogonalisation in FP64) down to the FP32 criterion, then
a short FP64 Lanczos restarted from that Ritz vector.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .device import stream_handle

FLOAT_CVG_FLOOR = 1.0e-7  # Lanczos_Gnd_Ut.cpp:185-193 caps CvgCrit for float operators


class LinOp:
    """Linear operator plug-in (include/LinOp.hpp:68-167): subclass and override ``matvec``."""

    def __init__(self, type: str = "mv", nx: int = 0, dtype: int = 3, device: int = -1):
        if type != "mv":
            raise ValueError("only the 'mv' LinOp form is supported")
        self._nx = int(nx)
        self._dtype = int(dtype)
        self._device = int(device)

    def nx(self) -> int:
        return self._nx

    def dtype(self) -> int:
        return self._dtype

    def device(self) -> int:
        return self._device

    def matvec(self, v):
        raise NotImplementedError("LinOp.matvec must be overridden")


class _DevView:
    """Exposes a raw HIP device pointer to torch (array interface, float64 vector)."""

    def __init__(self, ptr: int, n: int):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": "<f8", "data": (ptr, False),
                                         "version": 2, "strides": None}


_CTX_CACHE = {}
_MATVEC_T = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p)


def _context(n: int, maxiter: int, device: int):
    key = (n, maxiter, device)
    ctx = _CTX_CACHE.get(key)
    if ctx is None:
        h = ctypes.c_void_p(0)
        torch.cuda.set_device(device)
        _lib.check(_lib.lib().cxa_lanczos_create(n, maxiter, ctypes.byref(h)))
        ctx = h.value
        if len(_CTX_CACHE) > 8:
            for k, v in list(_CTX_CACHE.items()):
                _lib.lib().cxa_lanczos_destroy(v)
            _CTX_CACHE.clear()
        _CTX_CACHE[key] = ctx
    return ctx


def _krylov_cap(n: int, maxiter: int, budget_bytes: int = 2 << 30) -> int:
    return int(max(2, min(maxiter, budget_bytes // (8 * max(n, 1)) - 1)))


def lanczos_gnd_native(op, psi0: torch.Tensor, CvgCrit: float = 1e-14, Maxiter: int = 10000,
                       batch: int = 8, reorth: bool = False, graph: bool = False,
                       fn_ctx=None) -> Tuple[float, torch.Tensor, int, np.ndarray, np.ndarray]:
    """Run the C driver on a flat FP64 device vector.  Returns (E, evec, niter, alphas, betas).
    ``op`` is a planned Heff (native chain) unless ``fn_ctx`` gives (fnptr, ctx)."""
    if psi0.dtype != torch.float64 or psi0.device.type != "cuda":
        raise TypeError("Lanczos runs on FP64 device vectors")
    if CvgCrit <= 0:
        raise ValueError("[ERROR][Lanczos] converge criteria must >0")
    if Maxiter < 2:
        raise ValueError("[ERROR][Lanczos] Maxiter must >1")
    n = psi0.numel()
    dev = psi0.device.index
    cap = _krylov_cap(n, Maxiter)
    ctx = _context(n, cap, dev)
    fn, mctx = fn_ctx if fn_ctx is not None else op.native_matvec()
    E = ctypes.c_double(0)
    niter = ctypes.c_int(0)
    alphas = np.zeros(cap + 1)
    betas = np.zeros(cap + 1)
    evec = torch.empty_like(psi0)
    flags = (1 if reorth else 0) | (2 if graph else 0)
    _lib.check(_lib.lib().cxa_lanczos_run(
        ctx, fn, mctx, psi0.data_ptr(), float(CvgCrit), int(min(Maxiter, cap)), int(batch), flags,
        ctypes.byref(E), evec.data_ptr(), ctypes.byref(niter),
        alphas.ctypes.data_as(_lib.dblp), betas.ctypes.data_as(_lib.dblp), stream_handle(dev)))
    k = niter.value
    return E.value, evec, k, alphas[:k].copy(), betas[:k].copy()


def _python_matvec(Hop, template):
    n = template._storage.numel()

    def cb(_ctx, xp, yp, _stream):
        try:
            x = torch.as_tensor(_DevView(xp, n), device="cuda")
            y = torch.as_tensor(_DevView(yp, n), device="cuda")
            v = template.clone_meta()
            v._storage = x
            out = Hop.matvec(v)
            y.copy_(out._storage.reshape(-1))
            return 0
        except Exception as e:  # reported as a failed matvec
            _python_matvec.last_error = e
            return _lib.ERR_INVALID

    return _MATVEC_T(cb)


def Lanczos(Hop: LinOp, Tin, method: str = "Gnd", CvgCrit: float = 1.0e-14,
            Maxiter: int = 10000, k: int = 1, is_V: bool = True, is_row: bool = True,
            max_krydim: int = 0, verbose: bool = False):
    """linalg::Lanczos for UniTensors, method "Gnd" (returns [energy, eigvector])."""
    if method != "Gnd":
        raise NotImplementedError("only method='Gnd' (Lanczos_Gnd_Ut) is on the hot path")
    if Tin.dtype() != 3:
        raise TypeError("the GPU Lanczos runs real FP64 operators")
    cvg = float(CvgCrit)
    x0 = Tin._storage.reshape(-1)
    from .heff import _HeffBase
    if isinstance(Hop, _HeffBase):
        E, v, k_, _, _ = lanczos_gnd_native(Hop, x0, cvg, Maxiter)
    else:
        cb = _python_matvec(Hop, Tin)
        fn = ctypes.cast(cb, ctypes.c_void_p).value
        try:
            E, v, k_, _, _ = lanczos_gnd_native(None, x0, cvg, Maxiter, batch=1, fn_ctx=(fn, None))
        except _lib.CxaError as err:
            inner = getattr(_python_matvec, "last_error", None)
            if inner is not None:
                _python_matvec.last_error = None
                raise inner from err
            raise
    from .unitensor import _scalar_block
    e = _scalar_block(3, x0.device.index)
    e._storage.fill_(E)
    out = [e]
    if is_V:
        ev = Tin.clone_meta()
        ev._storage = v
        out.append(ev)
    return out


def ground_state_mixed(op, psi0: torch.Tensor, CvgCrit: float = 1e-12, Maxiter: int = 1000,
                       batch: int = 8, graph: bool = False, fp32_cvg: float = FLOAT_CVG_FLOOR):
    """FP32 + refinement: (E, evec, n_fp32_iter, n_fp64_iter).

    Phase 1 runs the contraction bulk on FP32 MFMA with FP64 vectors and FP64 full
    re-orthogonalisation until |dE| < fp32_cvg; phase 2 restarts an FP64 Lanczos from the phase-1
    Ritz vector with the caller's CvgCrit.  The eigensolve, norms and orthogonalisation are FP64
    throughout."""
    from .unitensor import PREC_FP32, PREC_FP64
    prev = op.precision
    try:
        op.set_precision(PREC_FP32)
        _, v1, n1, _, _ = lanczos_gnd_native(op, psi0, max(fp32_cvg, CvgCrit), Maxiter, batch,
                                             reorth=True, graph=graph)
        op.set_precision(PREC_FP64)
        E, v2, n2, _, _ = lanczos_gnd_native(op, v1, CvgCrit, Maxiter, batch, reorth=True,
                                             graph=graph)
    finally:
        op.set_precision(prev)
    return E, v2, n1, n2


__all__ = ["LinOp", "Lanczos", "lanczos_gnd_native", "ground_state_mixed"]
