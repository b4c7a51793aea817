"""Two-site U(1) DMRG on the GPU: the sweep of the reference's example
(example/DMRG/dmrg_two_sites_U1.py:44-205) with this package's operators.

Index conventions (those of the example and of ``heff.Heff2``):

    A_k    (a_k KET, s_k BRA, a_{k+1} BRA)       q(a_{k+1}) = q(a_k) - q(s_k), q(a_0) = 0
    L_k    (w BRA, a BRA, a' KET)                 left environment of site k
    R_k    (w KET, a KET, a' BRA)                 right environment of bond k-1 | k
    M      (w KET, w' BRA, s KET, s' BRA)         the XXZ MPO (heisenberg.xxz_mpo_block)

One bond update (p, p+1): psi = A_p . A_{p+1}; ground state of the two-site effective
Hamiltonian (Heff2: the L.psi.W12 mix launch and the .R grouped GEMM) by the native Lanczos
driver, FP64 or FP32 + FP64 refinement (in mixed mode a bond whose predecessor converged in
fewer than MIXED_MIN_ITERS iterations runs FP64 directly; operators of 2^17 elements or more
run tilings tuned once per size class and precision); Svd_truncate(psi, min(d dim_l, d dim_r,
chi)) with the
singular values normalised; the singular values absorbed towards the sweep direction; and the
environment of the site left behind updated by three block contractions
(L' = ((L.A).M).A^dagger, as the example's network "L: -2,-1,-3 / A: -1,-4,1 / M: -2,0,-4,-5 /
A_Conj: -3,-5,2").  A^dagger with the example's ``Dagger().permute_(labels)`` is the complex
conjugate with every bond redirected, in the same block layout.

The initial state is a random MPS at the target bond dimension (sectors from
``mps.bond_sectors``), brought to right-canonical form by Gesvd from the right; the sweeps then
run left to right and right to left.  Energies are recorded per bond update like the example's
``Ekeep``.
"""
from __future__ import annotations

import time
from typing import Dict, List, Optional

import numpy as np
import torch

from .heff import Heff2
from .heisenberg import MPO_GROUPS, mpo_bond, phys_bond, xxz_mpo_block
from .lanczos import ground_state_mixed, lanczos_gnd_native
from .mps import bond_sectors
from .svd import Gesvd, Svd_truncate, orthogonal_split, truncated_split
from .unitensor import BD_BRA, BD_KET, PREC_FP32, PREC_FP64, Bond, Contract, UniTensor


def _bra(A: UniTensor) -> UniTensor:
    """A.Dagger().permute_(A.labels()): conjugate data, every bond redirected, same layout."""
    out = A.Conj()
    out._bonds = [b.redirect() for b in out._bonds]
    return out


def left_env_step(Lk: UniTensor, A: UniTensor, M: UniTensor) -> UniTensor:
    """L_{k+1}(w', a_{k+1}, a'_{k+1}) from L_k(w, a, a'), A_k(a, s, a_{k+1}), M(w, w', s, t)."""
    L = Lk.relabel(["~w", "~a", "~b"])
    X = Contract(L, A.relabel(["~a", "~s", "~an"]))
    Y = Contract(X, M.relabel(["~w", "~wn", "~s", "~t"]))
    Z = Contract(Y, _bra(A).relabel(["~b", "~t", "~bn"]))
    return Z.permute(["~wn", "~an", "~bn"], 1)


def right_env_step(Rk: UniTensor, B: UniTensor, M: UniTensor) -> UniTensor:
    """R_k(w, a_k, a'_k) from R_{k+1}(w', a_{k+1}, a'_{k+1}), B_k(a_k, s, a_{k+1}),
    M(w, w', s, t)."""
    R = Rk.relabel(["~w", "~a", "~b"])
    X = Contract(R, B.relabel(["~an", "~s", "~a"]))
    Y = Contract(X, M.relabel(["~wn", "~w", "~s", "~t"]))
    Z = Contract(Y, _bra(B).relabel(["~bn", "~t", "~b"]))
    return Z.permute(["~wn", "~an", "~bn"], 1)


def _boundary(left: bool, device: int, q: int = 0) -> UniTensor:
    """L_0 or R_L on the 1-dimensional virtual bond of charge q: 1 at the MPO's left-boundary
    state (4, the identity row) or right-boundary state (0)."""
    vb = Bond(BD_KET, [[q]], [1])
    bw = mpo_bond()
    if left:
        E = UniTensor([bw.redirect(), vb.redirect(), vb], ["w", "a", "b"], 1, 3, -1)
        state = 4
    else:
        E = UniTensor([bw, vb, vb.redirect()], ["w", "a", "b"], 1, 3, -1)
        state = 0
    for bi, t in enumerate(E.itoi()):
        states = MPO_GROUPS[t[0]][1]
        if state in states:
            blk = np.zeros([int(s) for s in E._shapes[bi]])
            blk[states.index(state), 0, 0] = 1.0
            E.put_block(blk, bi)
    return E.to(device)


def random_mps(L: int, chi: int, seed: int = 1, device: int = 0, Q: int = 0) -> List[UniTensor]:
    """A_0..A_{L-1} (conventions above) with iid N(0,1) blocks at the bond dimensions of
    mps.bond_sectors in the sector of total charge Q = sum of 2Sz (bond charges negated to this
    convention: q(a_k) = -(charge of sites 0..k-1), so q(a_L) = -Q)."""
    rng = np.random.default_rng(seed)
    phys = phys_bond().redirect()
    bonds = []
    for k in range(L + 1):
        q, d = bond_sectors(L, k, chi, Q)
        bonds.append(Bond(BD_KET, [[-x] for x in q], d))
    As = []
    for k in range(L):
        A = UniTensor([bonds[k], phys, bonds[k + 1].redirect()],
                      ["a%d" % k, "s%d" % k, "a%d" % (k + 1)], 2, 3, -1, _alloc=False)
        n = int(sum(np.prod(s.astype(np.int64)) for s in A._shapes))
        A._storage = torch.from_numpy(rng.standard_normal(n)).to("cuda:%d" % device)
        As.append(A)
    return As


def _orthonormal_rows(Mb: torch.Tensor, rows: torch.Tensor, tol: float = 1e-14,
                      maxit: int = 80) -> torch.Tensor:
    """The polar factor of each full-row-rank matrix of the batch Mb (B x mp x np, zero padded;
    rows[b, i] = 1 on its true rows): X with orthonormal rows spanning the same row space.
    Newton-Schulz iteration X <- 1.5 X - 0.5 (X X^T) X from X = M / |M|_F (every singular
    value in (0, 1]), which maps each singular value s to 1.5 s - 0.5 s^3 (to 1, quadratically
    near it) and keeps the zero padding; only batched GEMMs.  Stops once max |X X^T - I| over
    the true rows is below tol; a batch that has not converged after maxit steps (a nearly
    rank-deficient matrix) is finished with a QR of X^T instead."""
    X = Mb / torch.linalg.vector_norm(Mb, dim=(1, 2), keepdim=True)
    eye = torch.diag_embed(rows)
    for it in range(maxit):
        XXt = torch.bmm(X, X.transpose(1, 2).contiguous())
        if it % 4 == 0 and float((XXt - eye).abs().max()) <= tol:
            return X
        X = 1.5 * X - 0.5 * torch.bmm(XXt, X)
    XXt = torch.bmm(X, X.transpose(1, 2).contiguous())
    if float((XXt - eye).abs().max()) <= tol:
        return X
    Q1, _ = torch.linalg.qr(X.transpose(1, 2).contiguous(), mode="reduced")
    return Q1.transpose(1, 2) * rows[:, :, None]


def random_right_canonical_mps(L: int, chi: int, seed: int = 1, device: int = 0,
                               Q: int = 0) -> List[UniTensor]:
    """A random right-canonical MPS (A_1..A_{L-1} right isometries, A_0 of unit norm) at the
    bond dimensions of mps.bond_sectors, reduced from the right to what each site can hold
    (a sector of a_k keeps at most as many states as its rows can be orthonormal: the columns
    (s, a_{k+1}) reaching it) -- the bond structure a right-to-left LQ sweep of random_mps
    produces.  Each site's sector matrices (rows a_k, columns (s, a_{k+1})) are iid N(0,1)
    orthonormalised (_orthonormal_rows); the sites are independent, so the sector matrices of
    all sites go through one batched iteration per padded size class."""
    from .svd import _pad_size, _sector_matrices
    rng = np.random.default_rng(seed)
    phys = phys_bond().redirect()
    target = []
    for k in range(L + 1):
        q, d = bond_sectors(L, k, chi, Q)
        target.append(Bond(BD_KET, [[-x] for x in q], d))
    bonds: List[Optional[Bond]] = [None] * (L + 1)
    bonds[0], bonds[L] = target[0], target[L]
    for k in range(L - 1, 0, -1):
        T = UniTensor([target[k], phys, bonds[k + 1].redirect()], ["a", "s", "b"], 1, 3, -1,
                      _alloc=False)
        avail = np.zeros(len(target[k].qnums()), dtype=np.int64)
        for t, shp in zip(T._itoi, T._shapes):
            avail[int(t[0])] += int(shp[1]) * int(shp[2])
        degs = np.minimum(np.asarray(target[k].getDegeneracies(), dtype=np.int64), avail)
        qs = [q for q, d in zip(target[k].qnums(), degs) if d > 0]
        bonds[k] = Bond(BD_KET, [[q] for q in qs], [int(d) for d in degs if d > 0])
    As = []
    for k in range(L):
        A = UniTensor([bonds[k], phys, bonds[k + 1].redirect()],
                      ["a%d" % k, "s%d" % k, "a%d" % (k + 1)], 1, 3, -1, _alloc=False)
        n = int(sum(np.prod(s.astype(np.int64)) for s in A._shapes))
        A._storage = torch.from_numpy(rng.standard_normal(n)).to("cuda:%d" % device)
        As.append(A)
    # every sector matrix of A_1..A_{L-1}: rows a_k (m), columns (s, a_{k+1}) (n >= m)
    work = []
    for k in range(1, L):
        for sec in _sector_matrices(As[k]):
            m, n = sec.m.shape
            if m > n:
                raise RuntimeError("[Internal ERROR] a sector has more rows than columns")
            work.append((k, sec))
    buckets: Dict[tuple, list] = {}
    for i, (_, sec) in enumerate(work):
        m, n = sec.m.shape
        buckets.setdefault((_pad_size(n), _pad_size(m)), []).append(i)
    for (np_, mp), idx in sorted(buckets.items()):
        Mb = torch.zeros((len(idx), mp, np_), dtype=torch.float64, device="cuda:%d" % device)
        rows = torch.zeros((len(idx), mp), dtype=torch.float64, device=Mb.device)
        for b, i in enumerate(idx):
            m, n = work[i][1].m.shape
            Mb[b, :m, :n] = work[i][1].m
            rows[b, :m] = 1.0
        Xb = _orthonormal_rows(Mb, rows)
        for b, i in enumerate(idx):
            k, sec = work[i]
            m, n = sec.m.shape
            sec.m = Xb[b, :m, :n]  # m x n, orthonormal rows
    # write the orthonormal rows back into the same blocks
    for k, sec in work:
        A = As[k]
        coff = np.concatenate([[0], np.cumsum(sec.csz)]).astype(int)
        pos = {tuple(int(x) for x in t): bi for bi, t in enumerate(A._itoi)}
        for j, ct in enumerate(sec.cols):
            bi = pos[tuple(sec.rows[0]) + tuple(ct)]
            o = int(A._offsets[bi])
            size = int(np.prod(A._shapes[bi].astype(np.int64)))
            A._storage[o:o + size] = sec.m[:, coff[j]:coff[j + 1]].reshape(-1)
    for A in As:
        A.set_rowrank_(2)
    As[0] = As[0] / As[0].Norm()
    return As


def right_canonicalize(As: List[UniTensor], svd: str = "gram") -> List[UniTensor]:
    """SVD from the right: A_k <- vT (right-orthonormal), A_{k-1} <- A_{k-1}.U.S; A_0 then
    normalised.  ``svd`` as in DMRG."""
    As = list(As)
    for k in range(len(As) - 1, 0, -1):
        A = As[k].clone_meta()
        A.set_rowrank_(1)
        if svd == "gram":
            US, vT = orthogonal_split(A)                      # US: (a_k, _aux_R)
        else:
            S, U, vT = Gesvd(A)
            US = Contract(U, S)                               # (a_k, _aux_R)
        As[k] = vT.relabel(As[k].labels()).set_rowrank_(2)
        prev = Contract(As[k - 1], US.relabel([As[k].labels()[0], "~aux"]))
        As[k - 1] = prev.relabel(As[k - 1].labels()).set_rowrank_(2)
    As[0] = As[0] / As[0].Norm()
    return As


class DMRG:
    """Two-site DMRG for the open XXZ chain (see module docstring)."""

    # operators with fewer elements run the default tilings (tuning would cost more than it
    # saves); larger ones are tuned once per precision and power-of-two size class
    TUNE_MIN_ELEMS = 1 << 17
    # in mixed mode a bond runs FP32 + refinement when the previous bond update took at least
    # this many iterations, plain FP64 otherwise: the refinement's few FP64 iterations only pay
    # off against a long FP32 phase (a converged sweep needs a handful of iterations per bond)
    MIXED_MIN_ITERS = 16

    def __init__(self, L: int, chi: int, J: float = 1.0, Jz: float = 1.0, seed: int = 1,
                 device: int = 0, maxit: int = 30, cvg: float = 1e-14, prec: str = "fp64",
                 mix: bool = True, Q: int = 0, svd: str = "gram", tune: bool = True):
        if prec not in ("fp64", "mixed"):
            raise ValueError("prec is 'fp64' or 'mixed'")
        if svd not in ("gram", "gesvd"):
            raise ValueError("svd is 'gram' or 'gesvd'")
        if L < 2:
            raise ValueError("DMRG needs at least 2 sites")
        self.L, self.chi, self.device = L, chi, device
        self.maxit, self.cvg, self.prec, self.mix, self.svd = maxit, cvg, prec, mix, svd
        self.tune = tune
        self._tilings: Dict[tuple, list] = {}
        self._last_iters: Optional[int] = None
        self.M = xxz_mpo_block(J, Jz, device=device)
        self.As = random_right_canonical_mps(L, chi, seed, device, Q)
        self.LR: List[Optional[UniTensor]] = [None] * (L + 1)
        self.LR[0] = _boundary(True, device)
        self.LR[L] = _boundary(False, device, -Q)
        for k in range(L - 1, 0, -1):
            self.LR[k] = right_env_step(self.LR[k + 1], self.As[k], self.M)
        self.energies: List[float] = []
        self.stats: Dict[str, float] = {"heff": 0.0, "lanczos": 0.0, "svd": 0.0, "env": 0.0,
                                        "matvecs": 0, "flops": 0.0, "updates": 0,
                                        "fp64_updates": 0}

    def _tile(self, H: Heff2, precs) -> None:
        """Tilings of the operator's launches: tuned on the first operator of each size class
        and precision (Heff2.tune), reused for the later ones."""
        if not self.tune or not H.chain.steps or H.nx() < self.TUNE_MIN_ELEMS:
            return
        size_class = int(np.log2(H.nx()))
        for prec in precs:
            H.set_precision(prec)
            choice = self._tilings.get((prec, size_class))
            if choice is None:
                self._tilings[(prec, size_class)] = H.tune(5)
            else:
                H.set_tiles(choice)

    def _optimize(self, p: int) -> UniTensor:
        t0 = time.perf_counter()
        A, B = self.As[p], self.As[p + 1]
        psi = Contract(A, B)                                  # (a_p, s_p, s_{p+1}, a_{p+2})
        labels = psi.labels()
        H = Heff2(psi.relabel(["-1", "-2", "-3", "-4"]),
                  self.LR[p].relabel(["-5", "-1", "0"]),
                  self.M.relabel(["-5", "-6", "-2", "1"]),
                  self.M.relabel(["-6", "-7", "-3", "2"]),
                  self.LR[p + 2].relabel(["-7", "-4", "3"]), mix=self.mix)
        mixed = self.prec == "mixed" and (self._last_iters is None or
                                          self._last_iters >= self.MIXED_MIN_ITERS)
        self._tile(H, (PREC_FP32, PREC_FP64) if mixed else (PREC_FP64,))
        H.set_precision(PREC_FP64)
        torch.cuda.synchronize(self.device)
        t1 = time.perf_counter()
        x0 = psi._storage / torch.linalg.vector_norm(psi._storage)
        if not mixed:
            E, v, n, _, _ = lanczos_gnd_native(H, x0, self.cvg, self.maxit)
            self._last_iters = n
            self.stats["fp64_updates"] += 1
        else:
            # the bond's iteration budget goes to the FP32 phase; the FP64 refinement polishes
            # the Ritz vector with a sixth of it (at least four iterations)
            E, v, n1, n2 = ground_state_mixed(H, x0, self.cvg, self.maxit,
                                              refine_maxiter=max(4, self.maxit // 6))
            n = n1 + n2
            self._last_iters = n1
        torch.cuda.synchronize(self.device)
        t2 = time.perf_counter()
        out = psi.clone_meta()
        out._storage = v
        out._labels = labels
        self.energies.append(float(E))
        st = self.stats
        st["heff"] += t1 - t0
        st["lanczos"] += t2 - t1
        st["matvecs"] += n
        st["flops"] += n * H.ref_flops
        st["updates"] += 1
        return out

    def _split(self, p: int, psi: UniTensor, absorb_left: bool) -> None:
        t0 = time.perf_counter()
        A, B = self.As[p], self.As[p + 1]
        dim_l, dim_r = A.shape()[0], B.shape()[2]
        d = A.shape()[1]
        new_dim = min(dim_l * d, dim_r * d, self.chi)
        psi.set_rowrank_(2)
        if self.svd == "gram":
            s, X, Y = truncated_split(psi, new_dim, "left" if absorb_left else "right")
            nrm = s.Norm()
            if absorb_left:
                X = X / nrm
            else:
                Y = Y / nrm
        else:
            s, U, vT = Svd_truncate(psi, new_dim)
            s = s / s.Norm()
            X, Y = (Contract(U, s), vT) if absorb_left else (U, Contract(s, vT))
        self.As[p] = X.relabel(A.labels()).set_rowrank_(2)
        self.As[p + 1] = Y.relabel(B.labels()).set_rowrank_(2)
        torch.cuda.synchronize(self.device)
        t1 = time.perf_counter()
        if absorb_left:
            self.LR[p + 1] = right_env_step(self.LR[p + 2], self.As[p + 1], self.M)
        else:
            self.LR[p + 1] = left_env_step(self.LR[p], self.As[p], self.M)
        torch.cuda.synchronize(self.device)
        t2 = time.perf_counter()
        self.stats["svd"] += t1 - t0
        self.stats["env"] += t2 - t1

    def sweep_left_to_right(self) -> None:
        for p in range(self.L - 1):
            self._split(p, self._optimize(p), absorb_left=False)

    def sweep_right_to_left(self) -> None:
        for p in range(self.L - 2, -1, -1):
            self._split(p, self._optimize(p), absorb_left=True)

    def run(self, sweeps: int) -> List[float]:
        """`sweeps` full sweeps (left to right, then right to left); energies per update."""
        for _ in range(sweeps):
            self.sweep_left_to_right()
            self.sweep_right_to_left()
        return self.energies

    def bond_dims(self) -> List[int]:
        return [A.shape()[2] for A in self.As[:-1]]


__all__ = ["DMRG", "random_mps", "right_canonicalize", "left_env_step", "right_env_step"]
