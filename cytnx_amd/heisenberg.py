"""Synthetic DMRG workloads: the spin-1/2 XXZ MPO, Gaussian U(1) sector profiles and random
(hermitised) environments, in the index conventions of the reference's two-site DMRG example
(example/DMRG/dmrg_two_sites_U1.py:30-35):

    psi : (l KET, s1 BRA, s2 BRA, r BRA)           labels -1,-2,-3,-4
    L   : (w BRA, l BRA, l' KET)                    labels -5,-1, 0
    M1  : (w KET, w' BRA, s1 KET, s1' BRA)          labels -5,-6,-2, 1
    M2  : (w' KET, w'' BRA, s2 KET, s2' BRA)        labels -6,-7,-3, 2
    R   : (w'' KET, r KET, r' BRA)                  labels -7,-4, 3
    out : (l' KET, s1' BRA, s2' BRA, r' BRA)        TOUT: 0,1;2,3

The MPO is the lower-triangular XXZ form
    W = [[I,0,0,0,0],[S+,0,0,0,0],[S-,0,0,0,0],[Sz,0,0,0,0],[0,J/2 S-,J/2 S+,Jz Sz,I]]
with W[w, w', s, s'] = <s'|O_{w,w'}|s>, basis (up, down) carrying U(1) charges 2Sz = (+1, -1).
The MPO bond states (0..4) carry charges (0, +2, -2, 0, 0); the U(1) MPO bond groups them as
qnums (0, +2, -2) with degeneracies (3, 1, 1), charge-0 states in the order (0, 3, 4).
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np

from .unitensor import BD_BRA, BD_KET, Bond, UniTensor

PHYS_Q = [1, -1]
MPO_STATE_Q = [0, 2, -2, 0, 0]
MPO_GROUPS = [(0, [0, 3, 4]), (2, [1]), (-2, [2])]


def gaussian_sectors(chi: int, qmax: int = 20, width: float = 50.0) -> Tuple[List[int], List[int]]:
    """Charges q in [-qmax, qmax] with sizes proportional to exp(-q^2/width), rounded by
    largest remainder to sum to chi; empty sectors dropped."""
    q = np.arange(-qmax, qmax + 1)
    w = np.exp(-(q.astype(np.float64) ** 2) / width)
    ideal = chi * w / w.sum()
    base = np.floor(ideal).astype(np.int64)
    rem = int(chi - base.sum())
    order = np.argsort(-(ideal - base), kind="stable")
    base[order[:rem]] += 1
    keep = base > 0
    return q[keep].astype(int).tolist(), base[keep].astype(int).tolist()


def xxz_mpo_dense(J: float = 1.0, Jz: float = 1.0) -> np.ndarray:
    """W[w, w', s, s'] (5, 5, 2, 2) as described in the module docstring."""
    I = np.eye(2)
    Sp = np.array([[0.0, 1.0], [0.0, 0.0]])  # <0|S+|1> = 1 (raises down to up)
    Sm = Sp.T.copy()
    Sz = np.diag([0.5, -0.5])
    ops = {(0, 0): I, (1, 0): Sp, (2, 0): Sm, (3, 0): Sz,
           (4, 1): 0.5 * J * Sm, (4, 2): 0.5 * J * Sp, (4, 3): Jz * Sz, (4, 4): I}
    W = np.zeros((5, 5, 2, 2))
    for (a, b), O in ops.items():
        W[a, b] = O.T  # W[.., s, s'] = <s'|O|s>
    return W


def mpo_bond() -> Bond:
    return Bond(BD_KET, [[q] for q, _ in MPO_GROUPS], [len(s) for _, s in MPO_GROUPS])


def phys_bond() -> Bond:
    return Bond(BD_KET, [[q] for q in PHYS_Q], [1, 1])


def xxz_mpo_block(J: float = 1.0, Jz: float = 1.0, device: int = -1) -> UniTensor:
    """The XXZ MPO as a U(1) block tensor M(w KET, w' BRA, s KET, s' BRA)."""
    W = xxz_mpo_dense(J, Jz)
    bw, bp = mpo_bond(), phys_bond()
    M = UniTensor([bw, bw.redirect(), bp, bp.redirect()], rowrank=2, device=-1)
    for bi, t in enumerate(M.itoi()):
        gw, gw2, s, s2 = t
        rows = MPO_GROUPS[gw][1]
        cols = MPO_GROUPS[gw2][1]
        blk = W[np.ix_(rows, cols, [s], [s2])]
        M.put_block(blk, bi)
    return M.to(device) if device >= 0 else M


def mpo_dense_from_block(M: UniTensor) -> np.ndarray:
    """Undo the grouping: dense (5,5,2,2) in the original state order."""
    D = M.to_dense()  # grouped order
    order = [s for _, states in MPO_GROUPS for s in states]
    inv = np.argsort(order)
    return D[np.ix_(inv, inv, [0, 1], [0, 1])]


def two_site_bonds(chi: int):
    q, d = gaussian_sectors(chi)
    vl = Bond(BD_KET, [[x] for x in q], d)
    return vl, q, d


def random_block(bonds: Sequence[Bond], labels: Sequence[str], rowrank: int, rng,
                 dtype=np.float64) -> UniTensor:
    T = UniTensor(bonds, labels, rowrank, 3 if dtype == np.float64 else 1, device=-1)
    x = rng.standard_normal(T.storage().numel())
    if dtype == np.complex128:
        x = x + 1j * rng.standard_normal(T.storage().numel())
    import torch
    T._storage = torch.from_numpy(x.astype(dtype))
    return T


def two_site_problem(chi: int, seed: int = 1, hermitian: bool = True, dtype=np.float64):
    """(psi, L, M1, M2, R) for the two-site U(1) Heff at bond dimension chi, on the host.

    Blocks are iid N(0,1) (complex128: real and imaginary parts iid N(0,1); the MPO stays
    real).  With hermitian=True the environments are symmetrised so that the effective
    Hamiltonian is Hermitian: L[w, l, l'] <- L[w, l, l'] + conj(L[wbar, l', l]) where wbar maps
    the MPO state 1 <-> 2 (S+ <-> S-), the same for R; then Lanczos applies."""
    rng = np.random.default_rng(seed)
    vl, q, d = two_site_bonds(chi)
    bp = phys_bond()
    bw = mpo_bond()
    psi = random_block([vl, bp.redirect(), bp.redirect(), vl.redirect()],
                       ["-1", "-2", "-3", "-4"], 1, rng, dtype)
    L = random_block([bw.redirect(), vl.redirect(), vl], ["-5", "-1", "0"], 1, rng, dtype)
    R = random_block([bw, vl, vl.redirect()], ["-7", "-4", "3"], 2, rng, dtype)
    if hermitian:
        L = _hermitise_env(L)
        R = _hermitise_env(R)
    M1 = xxz_mpo_block()
    M1.set_labels(["-5", "-6", "-2", "1"])
    M2 = xxz_mpo_block()
    M2.set_labels(["-6", "-7", "-3", "2"])
    return psi, L, M1, M2, R


def _hermitise_env(E: UniTensor) -> UniTensor:
    """E[w, a, b] + conj(E[wbar, b, a]) on the dense form, written back block by block."""
    D = E.to_dense()
    order = [s for _, states in MPO_GROUPS for s in states]   # grouped position -> state
    pos = {s: i for i, s in enumerate(order)}
    swap = {0: 0, 1: 2, 2: 1, 3: 3, 4: 4}
    perm = [pos[swap[order[i]]] for i in range(5)]           # grouped index of wbar
    D2 = D + np.conj(np.transpose(D[perm], (0, 2, 1)))
    out = E.clone_meta()
    import torch
    out._storage = torch.zeros_like(E._storage)
    offs = [np.concatenate([[0], np.cumsum(b.getDegeneracies())[:-1]]).astype(int)
            for b in E._bonds]
    for bi, t in enumerate(E._itoi):
        shp = [int(s) for s in E._shapes[bi]]
        sl = tuple(slice(offs[a][int(qi)], offs[a][int(qi)] + shp[a]) for a, qi in enumerate(t))
        out.put_block(D2[sl], bi)
    return out
