"""LinOp and ground-state Lanczos (the reference's linalg::Lanczos(method="Gnd")).

``Lanczos(Hop, Tin, method="Gnd", CvgCrit, Maxiter)`` mirrors the Python binding of
src/linalg/Lanczos.cpp:135-179 / Lanczos_Gnd_Ut.cpp: it returns ``[energy, eigvector]``.
The iteration itself runs in the C++ driver (cxa_lanczos_run) on device vectors:

* for the planned effective Hamiltonians (cytnx_amd.heff) the matvec is the C-side chain, so an
  iteration has no Python in it;
* any other LinOp subclass is called back through Python per matvec, like the reference's
  PyLinOp trampoline (pybind/linop_py.cpp:22-41).

``ground_state_mixed`` is the FP32 + refinement scheme: Lanczos with the contraction bulk on
FP32 MFMA (vectors and scalars in FP64) to three decades above the caller's criterion, then a
short FP64 Lanczos restarted from that Ritz vector.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from .device import stream_handle



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
    """Exposes a raw HIP device pointer to torch (array interface, float64 or complex128
    vector)."""

    def __init__(self, ptr: int, n: int, typestr: str = "<f8"):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False),
                                         "version": 2, "strides": None}


_CTX_CACHE = {}  # (n, maxiter, device) -> context handle, as last shaped
_CTX_CAP = {}  # context handle -> (n, maxiter, device) it was created for
_MATVEC_T = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                             ctypes.c_void_p)


CTX_CACHE_BYTES = 16 << 30  # Krylov bases kept alive between runs (all devices)
CTX_HEADROOM = 1.25  # a new context's vectors hold this many times the requested length


def _context(n: int, maxiter: int, device: int):
    """A Lanczos context of n-element vectors holding maxiter + 1 basis vectors, reused across
    runs: a cached context of the device whose buffers hold the request is reshaped to it
    (cxa_lanczos_reshape; the smallest such), otherwise a new one is made with CTX_HEADROOM
    times the vector length, so that runs of slowly varying size (DMRG bond updates) share
    contexts instead of freeing and allocating per run.  Cached contexts are released (all of
    them) once their bases would pass CTX_CACHE_BYTES."""
    key = (n, maxiter, device)
    ctx = _CTX_CACHE.get(key)
    if ctx is not None:
        return ctx
    fits = [(cn * (cm + 1), h) for h, (cn, cm, cd) in _CTX_CAP.items()
            if cd == device and cn >= n and cm >= maxiter]
    if fits:
        ctx = min(fits)[1]
        _lib.check(_lib.lib().cxa_lanczos_reshape(ctx, n, maxiter))
        for k in [k for k, v in _CTX_CACHE.items() if v == ctx]:
            del _CTX_CACHE[k]
        _CTX_CACHE[key] = ctx
        return ctx
    # the headroom stays within the basis budget that sized maxiter
    cap_n = max(n, min(int(n * CTX_HEADROOM), KRYLOV_BUDGET_BYTES // (8 * (maxiter + 1))))
    need = 8 * cap_n * (maxiter + 1)
    held = sum(8 * cn * (cm + 1) for cn, cm, _ in _CTX_CAP.values())
    if _CTX_CAP and held + need > CTX_CACHE_BYTES:
        torch.cuda.synchronize()
        for h in _CTX_CAP:
            _lib.lib().cxa_lanczos_destroy(h)
        _CTX_CACHE.clear()
        _CTX_CAP.clear()
    h = ctypes.c_void_p(0)
    torch.cuda.set_device(device)
    _lib.check(_lib.lib().cxa_lanczos_create(cap_n, maxiter, ctypes.byref(h)))
    ctx = h.value
    _CTX_CAP[ctx] = (cap_n, maxiter, device)
    if cap_n != n:
        _lib.check(_lib.lib().cxa_lanczos_reshape(ctx, n, maxiter))
    _CTX_CACHE[key] = ctx
    return ctx


KRYLOV_BUDGET_BYTES = 8 << 30  # device memory one Krylov basis may take


def _krylov_cap(n: int, maxiter: int, budget_bytes: int = KRYLOV_BUDGET_BYTES) -> int:
    """Basis vectors a context holds: at most maxiter, within the memory budget."""
    return int(max(2, min(maxiter, budget_bytes // (8 * max(n, 1)) - 1)))


class KrylovCapWarning(RuntimeWarning):
    """A Lanczos run stopped at the memory cap of its Krylov basis before its criterion."""


class KrylovCapError(RuntimeError):
    """A Lanczos_Exp run needs more Krylov vectors than the memory budget allows."""


def _doubles(x: torch.Tensor, who: str) -> Tuple[torch.Tensor, bool]:
    """(the flat vector as FP64 doubles, complex?) for the C driver: a complex128 vector is
    handed over as its (re, im) pairs with the driver's complex flag."""
    if x.device.type != "cuda" or x.dtype not in (torch.float64, torch.complex128):
        raise TypeError("%s runs on FP64 or complex128 device vectors" % who)
    if x.dtype == torch.complex128:
        return torch.view_as_real(x.contiguous()).reshape(-1), True
    return x.contiguous().reshape(-1), False


CPLX_FLAG = 4  # the driver's complex-Hermitian flag (vectors of (re, im) pairs)


def lanczos_gnd_native(op, psi0: torch.Tensor, CvgCrit: float = 1e-14, Maxiter: int = 10000,
                       batch: int = 8, reorth: bool = False, graph: bool = False,
                       fn_ctx=None) -> Tuple[float, torch.Tensor, int, np.ndarray, np.ndarray]:
    """Run the C driver on a flat FP64 or complex128 (Hermitian operator) device vector.
    Returns (E, evec, niter, alphas, betas); evec has psi0's dtype.  ``op`` is a planned Heff
    (native chain) unless ``fn_ctx`` gives (fnptr, ctx)."""
    x, cplx = _doubles(psi0, "Lanczos")
    if CvgCrit <= 0:
        raise ValueError("[ERROR][Lanczos] converge criteria must >0")
    if Maxiter < 2:
        raise ValueError("[ERROR][Lanczos] Maxiter must >1")
    n = x.numel()
    dev = psi0.device.index
    cap = _krylov_cap(n, Maxiter)
    ctx = _context(n, cap, dev)
    fn, mctx = fn_ctx if fn_ctx is not None else op.native_matvec()
    E = ctypes.c_double(0)
    niter = ctypes.c_int(0)
    alphas = np.zeros(cap + 1)
    betas = np.zeros(cap + 1)
    evec = torch.empty_like(x)
    flags = (1 if reorth else 0) | (2 if graph else 0) | (CPLX_FLAG if cplx else 0)
    _lib.check(_lib.lib().cxa_lanczos_run(
        ctx, fn, mctx, x.data_ptr(), float(CvgCrit), int(min(Maxiter, cap)), int(batch), flags,
        ctypes.byref(E), evec.data_ptr(), ctypes.byref(niter),
        alphas.ctypes.data_as(_lib.dblp), betas.ctypes.data_as(_lib.dblp), stream_handle(dev)))
    k = niter.value
    if cap < Maxiter and k >= cap and betas[k - 1] != 0.0:
        from scipy.linalg import eigvalsh_tridiagonal
        prev = eigvalsh_tridiagonal(alphas[:k - 1], betas[:k - 2], select="i",
                                    select_range=(0, 0))[0] if k > 2 else alphas[0]
        if not abs(E.value - prev) < CvgCrit:
            import warnings
            warnings.warn("Lanczos stopped after %d iterations, the Krylov basis the device "
                          "memory budget allows (%d vectors of %d elements), with |dE| = %.3e "
                          "above CvgCrit %.3e" % (k, cap, n, abs(E.value - prev), CvgCrit),
                          KrylovCapWarning, stacklevel=2)
    if cplx:
        evec = torch.view_as_complex(evec.view(-1, 2))
    return E.value, evec, k, alphas[:k].copy(), betas[:k].copy()


def _flat_vector(T) -> Tuple[torch.Tensor, object]:
    """(flat FP64 device data, a contiguous template with the same metadata) of a UniTensor (block
    or dense) or a Tensor."""
    from .tensor import Tensor
    from .unitensor import UniTensor
    if isinstance(T, UniTensor):
        t = T.clone()
        t.contiguous_()
        flat = t.storage().reshape(-1)
    elif isinstance(T, Tensor):
        t = T.contiguous()
        flat = t.storage().reshape(-1)
    else:
        raise TypeError("Tin must be a UniTensor or a Tensor")
    if flat.dtype not in (torch.float64, torch.complex128) or flat.device.type != "cuda":
        raise TypeError("the GPU Lanczos runs FP64 or complex128 operators on device tensors")
    return flat, t


def _wrap_like(template, x: torch.Tensor):
    """A tensor with the template's metadata over the flat data x (no copy)."""
    from .tensor import Tensor
    from .unitensor import UniTensor
    if isinstance(template, UniTensor):
        v = template.clone_meta()
        if template._is_block:
            v._storage = x
        else:
            v._block = Tensor.from_storage(x, template._block.shape())
        return v
    return Tensor.from_storage(x, template.shape())


def _flat_result(out) -> torch.Tensor:
    from .tensor import Tensor
    if isinstance(out, Tensor):
        return out.contiguous().storage().reshape(-1)
    out.contiguous_()
    return out.storage().reshape(-1)


def _python_matvec(Hop, template):
    st = template.storage()
    n = int(st.numel())  # the template is contiguous
    ts = "<c16" if st.dtype == torch.complex128 else "<f8"

    def cb(_ctx, xp, yp, _stream):
        try:
            x = torch.as_tensor(_DevView(xp, n, ts), device="cuda")
            y = torch.as_tensor(_DevView(yp, n, ts), device="cuda")
            out = Hop.matvec(_wrap_like(template, x))
            y.copy_(_flat_result(out))
            return 0
        except Exception as e:  # reported as a failed matvec
            _python_matvec.last_error = e
            return _lib.ERR_INVALID

    return _MATVEC_T(cb)


def _matvec_of(Hop, template):
    """(fnptr, ctx) of the matvec, plus the Python callback object to keep alive."""
    from .heff import _HeffBase
    if isinstance(Hop, _HeffBase):
        return Hop.native_matvec(), None
    cb = _python_matvec(Hop, template)
    return (ctypes.cast(cb, ctypes.c_void_p).value, None), cb


def _reraise_matvec(err):
    inner = getattr(_python_matvec, "last_error", None)
    if inner is not None:
        _python_matvec.last_error = None
        raise inner from err
    raise err


_WHICH = {"SA": 0, "LA": 1, "LM": 2}


def eigsh_native(op, psi0: torch.Tensor, which: str = "SA", k: int = 1, tol: float = 0.0,
                 want_vectors: bool = True, fn_ctx=None, Maxiter: int = 10000, ncv: int = 0):
    """k extreme eigenpairs of a flat FP64 or complex128 (Hermitian) device operator
    (cxa_lanczos_eigsh: thick-restart Lanczos with ncv basis vectors, 0 = min(n, 2k + 10), for at
    most Maxiter restart cycles).  Returns (eigenvalues (numpy, reference order), eigenvectors
    (k x n device tensor of psi0's dtype, or None), matvecs applied, pairs within the
    tolerance).  ``op`` is a planned Heff unless ``fn_ctx`` gives (fnptr, ctx)."""
    x, cplx = _doubles(psi0, "Lanczos")
    n = x.numel()
    ne = psi0.numel()
    dev = psi0.device.index
    m = min(ne, int(ncv) if ncv else 2 * int(k) + 10)
    ctx = _context(n, max(2, 2 * m), dev)
    fn, mctx = fn_ctx if fn_ctx is not None else op.native_matvec()
    evals = np.zeros(k)
    evecs = torch.empty((k, n), dtype=torch.float64, device=psi0.device) if want_vectors else None
    niter = ctypes.c_int(0)
    nconv = ctypes.c_int(0)
    _lib.check(_lib.lib().cxa_lanczos_eigsh(
        ctx, fn, mctx, x.data_ptr(), _WHICH[which], int(k), float(tol), int(Maxiter), int(ncv),
        CPLX_FLAG if cplx else 0, evals.ctypes.data_as(_lib.dblp),
        evecs.data_ptr() if evecs is not None else None, ctypes.byref(niter),
        ctypes.byref(nconv), stream_handle(dev)))
    if cplx and evecs is not None:
        evecs = torch.view_as_complex(evecs.view(k, ne, 2))
    return evals, evecs, niter.value, nconv.value


def exp_native(op, psi0: torch.Tensor, tau, CvgCrit: float = 1e-14,
               Maxiter: int = 10000, fn_ctx=None):
    """exp(tau H) psi0 for a flat device operator (cxa_lanczos_exp): (out, Krylov vectors used).
    psi0 FP64 (real operator) or complex128 (Hermitian operator); tau real or complex.  out is
    complex128 when psi0 is, or when tau has an imaginary part.  The basis holds at most
    min(Maxiter, n) vectors within the device memory budget; a run that reaches that budget
    before the reference's stop test raises KrylovCapError."""
    x, cplx = _doubles(psi0, "Lanczos_Exp")
    tau = complex(tau)
    n = x.numel()
    ne = psi0.numel()
    dev = psi0.device.index
    want = max(2, min(int(Maxiter), ne))
    cap = _krylov_cap(n, want)
    ctx = _context(n, cap, dev)
    fn, mctx = fn_ctx if fn_ctx is not None else op.native_matvec()
    cout = cplx or tau.imag != 0.0
    out = torch.empty(2 * ne if cout else ne, dtype=torch.float64, device=psi0.device)
    niter = ctypes.c_int(0)
    _lib.check(_lib.lib().cxa_lanczos_exp(ctx, fn, mctx, x.data_ptr(), float(tau.real),
                                          float(tau.imag), float(CvgCrit),
                                          int(min(Maxiter, cap)), CPLX_FLAG if cplx else 0,
                                          out.data_ptr(), ctypes.byref(niter),
                                          stream_handle(dev)))
    if cap < want and niter.value >= cap:
        raise KrylovCapError("Lanczos_Exp: the Krylov basis reached %d vectors of %d elements "
                             "(the device memory budget, %d bytes) before the stop test; pass "
                             "a smaller Maxiter or a smaller |tau|"
                             % (cap, n, KRYLOV_BUDGET_BYTES))
    if cout:
        out = torch.view_as_complex(out.view(ne, 2))
    return out, niter.value


def Lanczos(Hop: LinOp, Tin, method: str = "Gnd", CvgCrit: Optional[float] = None,
            Maxiter: int = 10000, k: int = 1, is_V: bool = True, is_row: bool = True,
            max_krydim: int = 0, verbose: bool = False, which: Optional[str] = None,
            ncv: int = 0):
    """linalg::Lanczos.

    * ``method="Gnd"`` (src/linalg/Lanczos.cpp:135-179, Lanczos_Gnd_Ut.cpp): ``[energy,
      eigvector]``, CvgCrit default 1e-14.
    * ``which="SA" | "LA" | "LM"`` (src/linalg/Lanczos.cpp:453-583, the ARPACK overload,
      pybind/linalg_py.cpp:722-740): ``k`` extreme eigenpairs, CvgCrit default 0 (machine
      precision).  UniTensor input: ``[eigenvalues, eigvector_1, ..., eigvector_k]``; Tensor
      input: ``[eigenvalues, eigenvectors]`` (shape (n,) for k = 1, (k, n) otherwise).
    """
    if which is not None:
        return _lanczos_which(Hop, Tin, which, Maxiter, 0.0 if CvgCrit is None else CvgCrit, k,
                              is_V, ncv, verbose)
    if method != "Gnd":
        raise NotImplementedError("method='Gnd', or which='SA'/'LA'/'LM' for k eigenpairs")
    if Tin.dtype() not in (1, 3):
        raise TypeError("the GPU Lanczos runs FP64 (Double) or complex128 (ComplexDouble) "
                        "operators")
    cvg = 1.0e-14 if CvgCrit is None else float(CvgCrit)
    x0, tmpl = _flat_vector(Tin)
    fn_ctx, keep = _matvec_of(Hop, tmpl)
    try:
        if keep is None:
            E, v, k_, _, _ = lanczos_gnd_native(Hop, x0, cvg, Maxiter)
        else:
            E, v, k_, _, _ = lanczos_gnd_native(None, x0, cvg, Maxiter, batch=1, fn_ctx=fn_ctx)
    except _lib.CxaError as err:
        _reraise_matvec(err)
    from .unitensor import _scalar_block
    e = _scalar_block(3, x0.device.index)  # the energy is real for both dtypes
    e._storage.fill_(E)
    out = [e]
    if is_V:
        out.append(_wrap_like(tmpl, v))
    return out


def _lanczos_which(Hop, Tin, which, Maxiter, CvgCrit, k, is_V, ncv, verbose):
    """k extreme eigenpairs on the device (cxa_lanczos_eigsh)."""
    import warnings
    from .tensor import Tensor
    from .unitensor import Bond, UniTensor
    if which not in _WHICH:
        raise ValueError("[ERROR][Lanczos] 'which' should be 'LM', 'LA', 'SA'")
    x0, tmpl = _flat_vector(Tin)
    n = x0.numel()
    k = int(k)
    if k < 1:
        raise ValueError("[ERROR][Lanczos] k should be >0")
    if k > n:
        raise ValueError("[ERROR][Lanczos] k can only be up to total dimension of input vector D")
    if CvgCrit < 0:
        raise ValueError("[ERROR][Lanczos] cvg_crit should be >= 0")
    if ncv != 0 and (ncv < 2 + k or ncv > n):
        raise ValueError("[ERROR][Lanczos] ncv should be 2+k<=ncv<=nx")
    if Maxiter < 1:
        raise ValueError("[ERROR][Lanczos] Maxiter must be positive")
    fn_ctx, keep = _matvec_of(Hop, tmpl)
    try:
        evals, evecs, niter, nconv = eigsh_native(None, x0, which, k, CvgCrit, is_V,
                                                  fn_ctx=fn_ctx, Maxiter=Maxiter, ncv=ncv)
    except _lib.CxaError as err:
        _reraise_matvec(err)
    del keep
    if nconv < k and verbose:
        warnings.warn("[WARNING][Lanczos]: %d of %d eigenpairs within the tolerance after %d "
                      "Krylov vectors" % (nconv, k, niter))
    dev = x0.device.index
    vals = torch.tensor(evals, dtype=torch.float64, device=x0.device)
    if evecs is not None:
        evecs = evecs.reshape(k, -1)
    if isinstance(Tin, UniTensor):
        ev = UniTensor([Bond(k)], rowrank=1, dtype=3, device=dev)
        ev._block = Tensor.from_storage(vals, [k])
        out = [ev]
        if is_V:
            out += [_wrap_like(tmpl, evecs[i].clone()) for i in range(k)]
        return out
    out = [Tensor.from_storage(vals, [k])]
    if is_V:
        out.append(Tensor.from_storage(evecs.reshape(-1), [n] if k == 1 else [k, n]))
    return out


def Lanczos_Exp(Hop: LinOp, v, tau, CvgCrit: float = 1.0e-14, Maxiter: int = 10000,
                verbose: bool = False):
    """exp(tau H) v by Lanczos (src/linalg/Lanczos_Exp.cpp:266-450) on the device
    (cxa_lanczos_exp): FP64 or complex128 (Hermitian) operators, real or complex tau (e.g.
    tau = -i dt for a real-time step).  Returns a tensor with v's metadata; its dtype is complex
    when v is complex or tau has an imaginary part (the reference's output dtype rule,
    Lanczos_Exp.cpp:67-76, 383-384)."""
    tau = complex(tau)
    if CvgCrit <= 0:
        raise ValueError("[ERROR][Lanczos_Exp] converge criteria must >0")
    if Maxiter < 2:
        raise ValueError("[ERROR][Lanczos_Exp] Maxiter must >1")
    x0, tmpl = _flat_vector(v)
    fn_ctx, keep = _matvec_of(Hop, tmpl)
    try:
        out, niter = exp_native(None, x0, tau, CvgCrit, Maxiter, fn_ctx=fn_ctx)
    except _lib.CxaError as err:
        _reraise_matvec(err)
    del keep
    if verbose:
        print("Lanczos_Exp: %d Krylov vectors" % niter)
    return _wrap_like(tmpl, out)  # the dtype follows out's storage (complex for complex tau)


def ground_state_mixed(op, psi0: torch.Tensor, CvgCrit: float = 1e-12, Maxiter: int = 1000,
                       batch: int = 16, graph: bool = False, fp32_cvg: Optional[float] = None,
                       reorth: bool = False, refine_maxiter: Optional[int] = None):
    """FP32 + refinement: (E, evec, n_fp32_iter, n_fp64_iter).

    Phase 1 is the Lanczos iteration with the contraction bulk on FP32 MFMA (the vectors, the
    recurrence scalars, the tridiagonal eigensolve and the Ritz vector in FP64) until
    |dE| < fp32_cvg, by default 1000 CvgCrit: the Ritz vector of an FP32 operator keeps
    converging after its energy has stopped improving (that energy carries the FP32 operator's
    ~1e-7 relative error), so the switch comes three decades above the target.  Phase 2 is the
    plain FP64 Lanczos restarted from that Ritz vector, stopped by the reference's rule
    |dE| < CvgCrit (Lanczos_Gnd_Ut.cpp:140-162); from a good start vector it needs a few
    iterations.  ``reorth`` applies to both phases (the reference's driver has none);
    ``refine_maxiter`` caps phase 2 (default Maxiter), for callers with a fixed iteration
    budget such as a DMRG sweep's bond updates."""
    from .unitensor import PREC_FP32, PREC_FP64
    if fp32_cvg is None:
        fp32_cvg = 1000.0 * CvgCrit
    prev = op.precision
    try:
        op.set_precision(PREC_FP32)
        _, v1, n1, _, _ = lanczos_gnd_native(op, psi0, max(fp32_cvg, CvgCrit), Maxiter, batch,
                                             reorth=reorth, graph=graph)
        op.set_precision(PREC_FP64)
        E, v2, n2, _, _ = lanczos_gnd_native(op, v1, CvgCrit,
                                             Maxiter if refine_maxiter is None
                                             else max(2, int(refine_maxiter)),
                                             batch, reorth=reorth, graph=graph)
    finally:
        op.set_precision(prev)
    return E, v2, n1, n2


__all__ = ["LinOp", "Lanczos", "Lanczos_Exp", "lanczos_gnd_native", "eigsh_native",
           "exp_native", "ground_state_mixed"]
