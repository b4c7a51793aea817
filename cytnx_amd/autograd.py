"""Reverse-mode differentiation of block-sparse (U(1)) contractions.

``contract_ad(A, B)`` is ``Contract(A, B)`` whose output storage carries a
``torch.autograd.Function`` node.  Its backward is a hand-written adjoint on the same grouped
MFMA GEMM: for C = A . B over the common legs,

    dA = dC . conj(B)   over B's free legs,  written into A's blocks with axes (free ++ common),
    dB = conj(A) . dC   over A's free legs,  written into B's blocks with axes (common ++ free),

each one planned contraction (the operands are read through strides, conjugated on load for
complex dtypes: ``cxa_block_contract_exec_ex``) followed by the inverse permute of every block
into the caller's axis order.  The adjoint outputs use A's (B's) own block list and offsets, so
the inverse permute lands directly in A's (B's) layout.  PyTorch's convention for complex
tensors holds: for a real loss, ``grad`` is the conjugate Wirtinger derivative, so
dLoss = Re(sum(conj(grad) * dA)).

``dagger_ad(A)`` is the reference's ``UniTensor.Dagger`` (conjugate, every bond redirected;
include/UniTensor.hpp:1707-1738) with the conjugation differentiated by PyTorch.
Cytnx itself has no automatic differentiation; this module is the differentiable front end the
MPS-gradient workload needs.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from . import _lib
from .device import stream_handle, torch_device
from .unitensor import (CPU, PREC_FP64, TORCH_OF, UniTensor, _common, _plan_for, _Plan,
                        _promote_pair, _scalar_block, block_permute_plan)


def _permuted_meta(T: UniTensor, axes: Sequence[int]) -> UniTensor:
    """T's blocks (same order, same offsets) with axes reordered to `axes`."""
    out = T.clone_meta()
    out._bonds = [T._bonds[a].clone() for a in axes]
    out._labels = [T._labels[a] for a in axes]
    out._itoi = np.ascontiguousarray(T._itoi[:, list(axes)])
    out._shapes = np.ascontiguousarray(T._shapes[:, list(axes)])
    out._storage = T._storage[:0]
    return out


def _block_permute(src: torch.Tensor, meta: UniTensor, mapper: Sequence[int],
                   dst: torch.Tensor) -> None:
    """Every block of `src` (layout `meta`) permuted by `mapper` into the same offsets of dst."""
    dt = {torch.float64: 3, torch.float32: 4, torch.complex128: 1, torch.complex64: 2}[src.dtype]
    plan = block_permute_plan(dt, src.device.index, meta._shapes, meta._offsets, mapper)
    _lib.check(_lib.lib().cxa_block_permute_exec(plan.handle, src.data_ptr(), dst.data_ptr(),
                                                 stream_handle(src.device.index)))


class _Spec:
    """Plans of one contraction structure: forward and the two adjoints."""

    def __init__(self, A: UniTensor, B: UniTensor, c1: List[int], c2: List[int],
                 out: UniTensor, dt: int):
        self.dt = dt
        self.out = out
        self.n_out = int(np.sum(np.prod(out._shapes.astype(np.int64), axis=1))) \
            if out.rank() and out.Nblocks() else max(1, out.Nblocks())
        self.fwd = _plan_for(dt, A, B, c1, c2, out)
        ra, rb = A.rank(), B.rank()
        nc1 = [i for i in range(ra) if i not in c1]
        nc2 = [i for i in range(rb) if i not in c2]
        na = len(nc1)
        # dA = dC (axes A free ++ B free) x B over B's free legs -> (A free ++ B common legs in
        # B's axis order); each B common leg stands for its A partner
        comm_by_b = sorted(range(len(c2)), key=lambda i: c2[i])
        pa = nc1 + [c1[i] for i in comm_by_b]
        self.Ap = _permuted_meta(A, pa)
        self.inv_a = [pa.index(j) for j in range(ra)]
        self.ga = _plan_for(dt, out, B, [na + k for k in range(len(nc2))], nc2, self.Ap) \
            if nc2 else None
        # dB = A x dC over A's free legs -> (A common legs in A's axis order ++ B free)
        comm_by_a = sorted(range(len(c1)), key=lambda i: c1[i])
        pb = [c2[i] for i in comm_by_a] + nc2
        self.Bp = _permuted_meta(B, pb)
        self.inv_b = [pb.index(j) for j in range(rb)]
        self.gb = _plan_for(dt, A, out, nc1, list(range(na)), self.Bp) if nc1 else None
        self.A, self.B = A, B

    def exec(self, plan: _Plan, L: torch.Tensor, R: torch.Tensor, C: torch.Tensor,
             flags: int) -> None:
        _lib.check(_lib.lib().cxa_block_contract_exec_ex(
            plan.handle, L.data_ptr(), R.data_ptr(), C.data_ptr(), plan.workspace(), PREC_FP64,
            flags, stream_handle(L.device.index)))


_SPECS: Dict[tuple, _Spec] = {}

# algorithmic flops issued by the forward and adjoint contractions (plans' sum of 2*M*N*K,
# x4 for complex), for throughput reports
FLOPS = {"forward": 0.0, "adjoint": 0.0}


def _spec_for(A: UniTensor, B: UniTensor, c1, c2, out: UniTensor, dt: int) -> _Spec:
    key = (dt, A.device(), A._itoi.tobytes(), A._shapes.tobytes(), B._itoi.tobytes(),
           B._shapes.tobytes(), tuple(c1), tuple(c2), out._itoi.tobytes(), A.rank(), B.rank())
    sp = _SPECS.get(key)
    if sp is None:
        if len(_SPECS) > 2048:
            _SPECS.clear()
        sp = _Spec(A, B, list(c1), list(c2), out, dt)
        _SPECS[key] = sp
    return sp


class BlockContractFn(torch.autograd.Function):
    """Storage-level autograd node of one block contraction (see module docstring)."""

    @staticmethod
    def forward(ctx, a: torch.Tensor, b: torch.Tensor, spec: _Spec) -> torch.Tensor:
        out = torch.empty(spec.n_out, dtype=a.dtype, device=a.device)
        spec.exec(spec.fwd, a, b, out, 0)
        FLOPS["forward"] += spec.fwd.flops
        ctx.save_for_backward(a, b)
        ctx.spec = spec
        return out

    @staticmethod
    def backward(ctx, g: torch.Tensor):
        a, b = ctx.saved_tensors
        sp: _Spec = ctx.spec
        g = g.contiguous()
        ga = gb = None
        if ctx.needs_input_grad[0]:
            tmp = torch.empty_like(a)
            if sp.ga is None:  # B has no free legs: dA = dC (scalar per A element) * conj(B)
                raise NotImplementedError("adjoint of a contraction that consumes all of B")
            sp.exec(sp.ga, g, b, tmp, 2)  # conj(B)
            FLOPS["adjoint"] += sp.ga.flops
            ga = torch.empty_like(a)
            _block_permute(tmp, sp.Ap, sp.inv_a, ga)
        if ctx.needs_input_grad[1]:
            tmp = torch.empty_like(b)
            if sp.gb is None:
                raise NotImplementedError("adjoint of a contraction that consumes all of A")
            sp.exec(sp.gb, a, g, tmp, 1)  # conj(A)
            FLOPS["adjoint"] += sp.gb.flops
            gb = torch.empty_like(b)
            _block_permute(tmp, sp.Bp, sp.inv_b, gb)
        return ga, gb, None


def contract_ad(A: UniTensor, B: UniTensor) -> UniTensor:
    """Contract(A, B) for U(1) block tensors on the GPU, differentiable in both storages
    (label bookkeeping, bond checks and output layout as Contract)."""
    if not (A.is_blockform() and B.is_blockform()):
        raise TypeError("contract_ad takes block UniTensors")
    comm, c1, c2 = _common(A.labels(), B.labels())
    for i, j in zip(c1, c2):
        if A._bonds[i].type() + B._bonds[j].type() != 0:
            raise ValueError("A bond with type BRA can only be contracted with type KET. "
                             "Invalid contraction for label '%s'." % A._labels[i])
    if A._storage.dtype != B._storage.dtype:
        raise TypeError("contract_ad: operands must share a dtype")
    dev = A.device()
    if dev == CPU or B.device() != dev:
        raise NotImplementedError("block contraction runs on one GPU; move both operands there")
    dt = A.dtype()
    nc1 = [i for i in range(A.rank()) if i not in c1]
    nc2 = [i for i in range(B.rank()) if i not in c2]
    out_bonds = [A._bonds[i] for i in nc1] + [B._bonds[i] for i in nc2]
    out_labels = [A._labels[i] for i in nc1] + [B._labels[i] for i in nc2]
    rr = A.rowrank() + B.rowrank() - sum(1 for i in c1 if i < A.rowrank()) \
        - sum(1 for i in c2 if i < B.rowrank())
    # the output's block list follows from the operands' layouts and the free bonds: reuse the
    # one made for the same structure (a sweep contracts the same structures again and again)
    key = (dt, dev, A._itoi.tobytes(), A._shapes.tobytes(), B._itoi.tobytes(),
           B._shapes.tobytes(), tuple(c1), tuple(c2),
           tuple((b._type, tuple(b._qnums), tuple(b._degs), b._dim) for b in out_bonds))
    hit = _OUTS.get(key)
    if hit is None:
        if out_bonds:
            tmpl = UniTensor(out_bonds, out_labels, max(0, min(rr, len(out_bonds))), dt, dev,
                             _alloc=False)
        else:
            tmpl = _scalar_block(dt, dev)
            tmpl._storage = tmpl._storage[:0]
        hit = (tmpl, _spec_for(A, B, c1, c2, tmpl, dt))
        if len(_OUTS) > 2048:
            _OUTS.clear()
        _OUTS[key] = hit
    tmpl, sp = hit
    out = tmpl.clone_meta()
    out._labels = out_labels
    if out_bonds:
        out._rowrank = max(0, min(rr, len(out_bonds)))
    out._storage = BlockContractFn.apply(A._storage, B._storage, sp)
    return out


_OUTS: Dict[tuple, Tuple[UniTensor, _Spec]] = {}


def dagger_ad(A: UniTensor) -> UniTensor:
    """Conjugate with every bond redirected (UniTensor::Dagger), differentiable."""
    out = A.clone_meta()
    out._bonds = [b.redirect() for b in A._bonds]
    st = A._storage
    out._storage = torch.conj_physical(st) if st.is_complex() else st
    return out


__all__ = ["contract_ad", "dagger_ad", "BlockContractFn"]
