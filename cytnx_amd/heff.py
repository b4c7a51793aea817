"""Effective-Hamiltonian matvecs for DMRG, planned once and applied as one C-side chain.

``Heff2`` is the two-site U(1) operator of the reference's DMRG example
(example/DMRG/dmrg_two_sites_U1.py:12-41), network

    psi: -1,-2,-3,-4 ; L: -5,-1,0 ; R: -7,-4,3 ; M1: -5,-6,-2,1 ; M2: -6,-7,-3,2 ; TOUT: 0,1;2,3

contracted in the fixed order L.psi -> M1 -> M2 -> R (the order of the reference's own
benchmark, benchmarks/UniTensor_bm.cpp:56, reversed in association: every step is one
block-sparse contraction).  The environments are permuted once at construction, L to
(l, w, l') and R to (r, w'', r'), so every block operand of every step is a strided matrix and
the grouped GEMMs read them in place -- the permutes of the reference's per-pair
``permute().reshape()`` (src/BlockUniTensor.cpp:1121-1133) disappear from the matvec.

``Heff1Dense`` is the dense one-site operator L.A.W.R of the C1 workload with the order
((L.A).W).R; its tensors are single blocks.

Both expose ``matvec`` (a cytnx_amd.lanczos.LinOp) and ``flops`` -- the algorithmic flop count
of one matvec summed over the block pairs each step pairs, as the reference's contract does.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from .device import stream_handle, torch_device
from .lanczos import LinOp, lanczos_gnd_native
from .unitensor import (BD_BRA, BD_KET, PREC_FP32, PREC_FP64, Bond, Contract, UniTensor,
                        contract_meta, meta_elems, plan_contraction)


class _Chain:
    def __init__(self, steps, device: int):
        """steps: list of (plan, fixed UniTensor, fixed_is_left, out meta UniTensor)."""
        self.device = device
        self.steps = steps
        self._plans = [s[0] for s in steps]  # keep the plans alive
        self._fixed = [s[1] for s in steps]
        n = len(steps)
        VP = ctypes.c_void_p * n
        h = ctypes.c_void_p(0)
        _lib.check(_lib.lib().cxa_chain_create(
            n, VP(*[p.handle for p in self._plans]),
            VP(*[f._storage.data_ptr() for f in self._fixed]),
            _lib.int_array([1 if s[2] else 0 for s in steps]),
            _lib.i64_array([f._storage.numel() for f in self._fixed]),
            _lib.i64_array([meta_elems(s[3]) for s in steps]), ctypes.byref(h)))
        self.handle = h.value
        f = ctypes.c_double(0)
        _lib.check(_lib.lib().cxa_chain_flops(self.handle, ctypes.byref(f)))
        self.flops = f.value
        self.prec = None

    def prepare(self, prec: int) -> None:
        if self.prec != prec:
            _lib.check(_lib.lib().cxa_chain_prepare(self.handle, prec,
                                                    stream_handle(self.device)))
            self.prec = prec

    def apply(self, x: torch.Tensor, y: torch.Tensor) -> None:
        _lib.check(_lib.lib().cxa_chain_matvec(self.handle, x.data_ptr(), y.data_ptr(),
                                               stream_handle(self.device)))

    def graph(self, x: torch.Tensor, y: torch.Tensor, repeats: int = 1) -> "ChainGraph":
        """The chain applied to fixed x -> y, captured once and replayed with one launch."""
        return ChainGraph(self, x, y, repeats)

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().cxa_chain_destroy(self.handle)
        except Exception:
            pass


class ChainGraph:
    """Captured chain application (cxa_chain_graph_create); keeps its vectors alive."""

    def __init__(self, chain: _Chain, x: torch.Tensor, y: torch.Tensor, repeats: int):
        self.chain, self.x, self.y, self.repeats = chain, x, y, repeats
        h = ctypes.c_void_p(0)
        torch.cuda.synchronize(chain.device)
        _lib.check(_lib.lib().cxa_chain_graph_create(chain.handle, x.data_ptr(), y.data_ptr(),
                                                     int(repeats), ctypes.byref(h)))
        self.handle = h.value

    def launch(self) -> None:
        _lib.check(_lib.lib().cxa_graph_launch(self.handle, stream_handle(self.chain.device)))

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().cxa_graph_destroy(self.handle)
        except Exception:
            pass


class _HeffBase(LinOp):
    """Shared driver of the planned chains: FP64 storage of the vector, selectable precision
    of the contraction bulk."""

    psi_template: UniTensor
    chain: _Chain

    def __init__(self, nx: int, device: int):
        super().__init__("mv", nx, 3, device)
        self.precision = PREC_FP64

    @property
    def flops(self) -> float:
        return self.chain.flops

    def set_precision(self, prec: int) -> None:
        self.chain.prepare(prec)
        self.precision = prec

    def apply_flat(self, x: torch.Tensor, y: torch.Tensor) -> None:
        """y = H x on flat FP64 device vectors in the psi block layout."""
        self.chain.prepare(self.precision)
        self.chain.apply(x, y)

    def matvec(self, v: UniTensor) -> UniTensor:
        out = v.clone_meta()
        out._storage = torch.empty_like(v._storage)
        self.apply_flat(v._storage, out._storage)
        return out

    def tune(self, reps: int = 20) -> List[dict]:
        """Autotune every step's tiling at the current precision on this operator's own
        buffers (cxa_chain_tune); returns the per-step choice and its time."""
        self.chain.prepare(self.precision)
        n = len(self.chain.steps)
        x = self.psi_template._storage if self.psi_template._storage.numel() else None
        if x is None:
            x = torch.zeros(self.nx(), dtype=torch.float64, device=torch_device(self.device()))
        y = torch.empty_like(x)
        kinds, splits = (ctypes.c_int * n)(), (ctypes.c_int * n)()
        ms = (ctypes.c_double * n)()
        _lib.check(_lib.lib().cxa_chain_tune(self.chain.handle, x.data_ptr(), y.data_ptr(),
                                             int(reps), kinds, splits, ms,
                                             stream_handle(self.device())))
        return [{"kind": kinds[i], "split": splits[i], "ms": ms[i]} for i in range(n)]

    def native_matvec(self):
        """(function pointer, context) for the C Lanczos driver."""
        self.chain.prepare(self.precision)
        fn = ctypes.cast(_lib.lib().cxa_chain_matvec, ctypes.c_void_p).value
        return fn, self.chain.handle

    def ground_state(self, psi0: torch.Tensor, CvgCrit: float = 1e-14, Maxiter: int = 10000,
                     batch: int = 8, reorth: bool = False, graph: bool = False):
        """Lanczos ground state on flat vectors (E, evec, niter)."""
        return lanczos_gnd_native(self, psi0, CvgCrit, Maxiter, batch, reorth, graph)


class Heff2(_HeffBase):
    """Two-site U(1) effective Hamiltonian (see module docstring).

    With ``fuse_mpo`` (default) the two MPO tensors are contracted once into the two-site MPO
    W12 = M1.M2 (over the shared MPO bond), and the matvec runs three grouped launches:
    L.psi, .W12, .R -- the intermediate of the M1 step is never stored.  ``flops`` is always the
    algorithmic count of the reference's four-step order L.psi -> M1 -> M2 -> R."""

    def __init__(self, psi: UniTensor, L: UniTensor, M1: UniTensor, M2: UniTensor,
                 R: UniTensor, fuse_mpo: bool = True):
        for T in (psi, L, M1, M2, R):
            if not T.is_blockform():
                raise TypeError("Heff2 takes U(1) block tensors")
            if T.dtype() != 3:
                raise TypeError("Heff2 runs real FP64 tensors")
        dev = L.device()
        if dev < 0:
            raise ValueError("move the environments to the GPU first")
        self.psi_template = psi.clone_meta()
        lab_w, lab_l, lab_lp = L.labels()    # L: (w, l, l')
        lab_w2, lab_r, lab_rp = R.labels()   # R: (w'', r, r')
        Lp = L.permute([lab_l, lab_w, lab_lp])   # (l, w, l')
        Rp = R.permute([lab_r, lab_w2, lab_rp])  # (r, w'', r')
        psi_m = psi.clone_meta()
        ref_steps = []
        X, c1, c2 = contract_meta(Lp, psi_m, 3, dev)
        ref_steps.append((plan_contraction(Lp, psi_m, X, c1, c2, 3), Lp, True, X))
        Y, c1, c2 = contract_meta(X, M1, 3, dev)
        ref_steps.append((plan_contraction(X, M1, Y, c1, c2, 3), M1, False, Y))
        Z, c1, c2 = contract_meta(Y, M2, 3, dev)
        ref_steps.append((plan_contraction(Y, M2, Z, c1, c2, 3), M2, False, Z))
        O, c1, c2 = contract_meta(Z, Rp, 3, dev)
        ref_steps.append((plan_contraction(Z, Rp, O, c1, c2, 3), Rp, False, O))
        if O.itoi() != psi.itoi() or [b.dim() for b in O.bonds()] != psi.shape():
            raise ValueError("Heff2: output block layout differs from psi's")
        self.ref_flops = float(sum(st[0].flops for st in ref_steps))
        self.pairs = [st[0].npairs for st in ref_steps]
        steps = ref_steps
        if fuse_mpo:
            W12 = Contract(M1, M2)
            Zf, c1, c2 = contract_meta(X, W12, 3, dev)
            if Zf.labels() != Z.labels() or Zf.itoi() != Z.itoi():
                raise ValueError("Heff2: fused MPO step changed the intermediate layout")
            steps = [ref_steps[0], (plan_contraction(X, W12, Zf, c1, c2, 3), W12, False, Zf),
                     ref_steps[3]]
        self.fused = bool(fuse_mpo)
        self.out_labels = O.labels()
        self.copies = [st[0].ws_bytes for st in steps]
        self.step_names = (["L.psi", ".W12", ".R"] if fuse_mpo
                           else ["L.psi", ".M1", ".M2", ".R"])
        self.chain = _Chain(steps, dev)
        super().__init__(meta_elems(psi), dev)

    @property
    def flops(self) -> float:
        return self.ref_flops


def _trivial_bond(dim: int, btype: int) -> Bond:
    return Bond(btype, [[0]], [dim])


def dense_as_block(a: torch.Tensor, labels: Sequence[str], types: Sequence[int]) -> UniTensor:
    """A dense array as a one-block U(1) tensor (all charges 0)."""
    bonds = [_trivial_bond(int(d), t) for d, t in zip(a.shape, types)]
    dev = -1 if a.device.type == "cpu" else a.device.index
    T = UniTensor(bonds, labels, 1, 3, dev, _alloc=False)
    T._storage = a.reshape(-1).contiguous().to(torch.float64)
    return T


class Heff1Dense(_HeffBase):
    """Dense one-site L(l,w,l') A(l,s,r) W(w,w',s,s') R(r,w',r') with order ((L.A).W).R."""

    def __init__(self, L: torch.Tensor, W: torch.Tensor, R: torch.Tensor, A_shape):
        dev = L.device.index
        Lt = dense_as_block(L, ["l", "w", "lp"], [BD_BRA, BD_BRA, BD_KET])
        Wt = dense_as_block(W, ["w", "wp", "s", "sp"], [BD_KET, BD_BRA, BD_KET, BD_BRA])
        Rt = dense_as_block(R, ["r", "wp", "rp"], [BD_KET, BD_KET, BD_BRA])
        At = UniTensor([_trivial_bond(int(d), t) for d, t in zip(A_shape, [BD_KET, BD_BRA,
                                                                          BD_BRA])],
                       ["l", "s", "r"], 1, 3, dev, _alloc=False)
        self.psi_template = At
        steps = []
        X, c1, c2 = contract_meta(Lt, At, 3, dev)
        steps.append((plan_contraction(Lt, At, X, c1, c2, 3), Lt, True, X))
        Y, c1, c2 = contract_meta(X, Wt, 3, dev)
        steps.append((plan_contraction(X, Wt, Y, c1, c2, 3), Wt, False, Y))
        O, c1, c2 = contract_meta(Y, Rt, 3, dev)
        steps.append((plan_contraction(Y, Rt, O, c1, c2, 3), Rt, False, O))
        if O.labels() != ["lp", "sp", "rp"]:
            raise ValueError("Heff1Dense: unexpected output order %s" % O.labels())
        self.pairs = [s[0].npairs for s in steps]
        self.copies = [s[0].ws_bytes for s in steps]
        self.chain = _Chain(steps, dev)
        super().__init__(int(np.prod(A_shape)), dev)


__all__ = ["Heff2", "Heff1Dense", "dense_as_block"]
