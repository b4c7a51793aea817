"""Differentiable energy of a finite U(1) MPS under the XXZ MPO.

Tensors and legs (U(1) charges are 2 Sz, KET bonds add +q, BRA bonds -q):

    A_k   (a_k BRA, s_k BRA, a_{k+1} KET)       site tensor, q(a_{k+1}) = q(a_k) + q(s_k)
    W_k   (w_k KET, w_{k+1} BRA, s_k KET, t_k BRA)   XXZ MPO, W[w, w', s, t] = <t|O_{w w'}|s>
    Ab_k  = Dagger(A_k) relabelled (b_k, t_k, b_{k+1})  bra tensor
    E_k   (a_k KET, w_k BRA, b_k BRA)           left environment

E_{k+1} = ((E_k . A_k) . W_k) . Ab_k (three block contractions), E_0 = the MPO's left boundary
state (the identity row, grouped index 2 of the charge-0 MPO sector) on 1-dimensional virtual
bonds, and <psi|H|psi> = E_L[0, right boundary state (grouped index 0), 0]; the norm uses the
same sweep without W.  Every contraction is ``contract_ad``, so torch autograd differentiates E
through the hand-written adjoints; with ``checkpoint`` each site's environment step is
recomputed in the backward pass instead of stored.

The bond structure of a random MPS: bond k carries the charges q = -k, -k+2, ..., k that can
still reach the total charge with the remaining sites, with degeneracies min(#left states,
#right states) of that charge, scaled down (largest remainder, at least 1 per charge) so that
their sum is at most chi.
"""
from __future__ import annotations

from math import comb
from typing import List, Sequence, Tuple

import numpy as np
import torch
import torch.utils.checkpoint

from .autograd import contract_ad, dagger_ad
from .heisenberg import MPO_GROUPS, xxz_mpo_block
from .unitensor import BD_BRA, BD_KET, Bond, UniTensor


def bond_sectors(L: int, k: int, chi: int, Q: int = 0) -> Tuple[List[int], List[int]]:
    """Charges and degeneracies of virtual bond k (0..L) of an L-site spin-1/2 MPS in the sector
    of total charge Q."""
    qs, full = [], []
    for q in range(-k, k + 1, 2):
        r = Q - q  # charge the remaining L-k sites carry
        if abs(r) > L - k or (L - k + r) % 2:
            continue
        left = comb(k, (k + q) // 2)
        right = comb(L - k, (L - k + r) // 2)
        qs.append(q)
        full.append(min(left, right))
    full = np.asarray(full, dtype=np.float64)
    total = full.sum()
    if total <= chi:
        return qs, full.astype(int).tolist()
    n = len(qs)
    ideal = np.maximum(1.0, full * chi / total) if chi >= n else full * chi / total
    base = np.floor(ideal).astype(np.int64)
    rem = int(chi - base.sum())
    if rem > 0:
        order = np.argsort(-(ideal - base), kind="stable")
        base[order[:rem]] += 1
    while rem < 0:
        # the floor of 1 per charge overshot chi: take the excess from the largest sectors
        order = np.argsort(-base, kind="stable")
        take = min(-rem, int(np.count_nonzero(base > 1)))
        if take == 0:
            break
        base[order[:take]] -= 1
        rem += take
    keep = base > 0
    return [q for q, k_ in zip(qs, keep) if k_], base[keep].astype(int).tolist()


def random_mps(L: int, chi: int, seed: int = 1, dtype=torch.complex128, device: int = 0,
               Q: int = 0) -> List[UniTensor]:
    """A_0..A_{L-1} with iid N(0,1) entries (complex: real and imaginary parts), each scaled by
    1/sqrt(2 dim(a_k)) so that norms stay O(1) along the chain."""
    rng = np.random.default_rng(seed)
    dt_id = {torch.complex128: 1, torch.complex64: 2, torch.float64: 3, torch.float32: 4}[dtype]
    phys = Bond(BD_BRA, [[1], [-1]], [1, 1])
    bonds = []
    for k in range(L + 1):
        q, d = bond_sectors(L, k, chi, Q)
        bonds.append((q, d))
    As = []
    for k in range(L):
        ql, dl = bonds[k]
        qr, dr = bonds[k + 1]
        bl = Bond(BD_BRA, [[x] for x in ql], dl)
        br = Bond(BD_KET, [[x] for x in qr], dr)
        A = UniTensor([bl, phys, br], ["a%d" % k, "s%d" % k, "a%d" % (k + 1)], 2, dt_id, -1,
                      _alloc=False)
        n = int(sum(np.prod(s.astype(np.int64)) for s in A._shapes))
        x = rng.standard_normal(n)
        if dtype.is_complex:
            x = x + 1j * rng.standard_normal(n)
        x = x / np.sqrt(2.0 * sum(dl))
        A._storage = torch.from_numpy(x).to(dtype).to("cuda:%d" % device)
        As.append(A)
    return As


def _mpo_site(W: UniTensor, k: int) -> UniTensor:
    return W.relabel(["w%d" % k, "w%d" % (k + 1), "s%d" % k, "t%d" % k])


def _boundary_env(A0: UniTensor, W: UniTensor, with_mpo: bool) -> UniTensor:
    """E_0 (a_0 KET, w_0 BRA, b_0 BRA) or F_0 (a_0 KET, b_0 BRA) on the 1-dimensional bond."""
    ba = A0.bonds()[0].redirect()          # KET
    bb = A0.bonds()[0]                     # BRA (the bra tensor's bond, redirected twice)
    dt = A0.dtype()
    dev = A0.device()
    if with_mpo:
        bw = W.bonds()[0].redirect()       # BRA
        E = UniTensor([ba, bw, bb], ["a0", "w0", "b0"], 1, dt, -1)
        for bi, t in enumerate(E.itoi()):
            blk = np.zeros([int(s) for s in E._shapes[bi]])
            if MPO_GROUPS[t[1]][0] == 0:
                blk[0, MPO_GROUPS[t[1]][1].index(4), 0] = 1.0  # identity row: left boundary
            E.put_block(blk, bi)
    else:
        E = UniTensor([ba, bb], ["a0", "b0"], 1, dt, -1)
        for bi in range(E.Nblocks()):
            E.put_block(np.ones([int(s) for s in E._shapes[bi]]), bi)
    return E.to(dev)


def _bra(A: UniTensor, k: int, phys_label: str) -> UniTensor:
    return dagger_ad(A).relabel(["b%d" % k, phys_label, "b%d" % (k + 1)])


def _env_step(E: UniTensor, A: UniTensor, W: UniTensor, k: int) -> UniTensor:
    X = contract_ad(E, A)                    # (w_k, b_k, s_k, a_{k+1})
    Y = contract_ad(X, _mpo_site(W, k))      # (b_k, a_{k+1}, w_{k+1}, t_k)
    return contract_ad(Y, _bra(A, k, "t%d" % k))   # (a_{k+1}, w_{k+1}, b_{k+1})


def _norm_step(F: UniTensor, A: UniTensor, k: int) -> UniTensor:
    X = contract_ad(F, A)                    # (b_k, s_k, a_{k+1})
    return contract_ad(X, _bra(A, k, "s%d" % k))   # (a_{k+1}, b_{k+1})


def _run(step, E0: UniTensor, As: Sequence[UniTensor], checkpoint: bool) -> UniTensor:
    E = E0
    for k, A in enumerate(As):
        if not checkpoint:
            E = step(E, A, k)
            continue
        meta_E, meta_A = E, A

        def fn(e_st, a_st, meta_E=meta_E, meta_A=meta_A, k=k):
            e = meta_E.clone_meta()
            e._storage = e_st
            a = meta_A.clone_meta()
            a._storage = a_st
            out = step(e, a, k)
            fn.meta = out
            return out._storage

        st = torch.utils.checkpoint.checkpoint(fn, E._storage, A._storage, use_reentrant=False)
        E = fn.meta.clone_meta()
        E._storage = st
    return E


def mps_energy(As: Sequence[UniTensor], W: UniTensor = None,
               checkpoint: bool = False) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(E, <psi|H|psi>, <psi|psi>) as differentiable torch scalars (real)."""
    if W is None:
        W = xxz_mpo_block(device=As[0].device())
        if As[0]._storage.is_complex():
            W = W.astype(As[0].dtype())
        elif As[0].dtype() != W.dtype():
            W = W.astype(As[0].dtype())
    L = len(As)
    E = _run(lambda e, a, k: _env_step(e, a, W, k), _boundary_env(As[0], W, True), As,
             checkpoint)
    F = _run(lambda f, a, k: _norm_step(f, a, k), _boundary_env(As[0], W, False), As,
             checkpoint)
    # E_L: (a_L KET, w_L BRA, b_L BRA), 1-dimensional virtual bonds; right boundary state =
    # grouped index 0 of the charge-0 MPO sector
    num = None
    for bi, t in enumerate(E.itoi()):
        if MPO_GROUPS[t[1]][0] == 0:
            num = E._storage[int(E._offsets[bi]) + MPO_GROUPS[t[1]][1].index(0)]
    den = F._storage.sum()
    num = num.real if num.is_complex() else num
    den = den.real if den.is_complex() else den
    return num / den, num, den


__all__ = ["bond_sectors", "random_mps", "mps_energy"]
