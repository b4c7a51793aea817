"""Singular value decompositions of UniTensors on the GPU: Gesvd / Svd / Svd_truncate.

Block (U(1)) tensors (Gesvd_BlockUT_internal, src/linalg/Gesvd.cpp:191-385):

* the rowrank splits the legs into rows and columns; a block's sector key is the charge its row
  legs carry, sum(+q for KET legs, -q for BRA legs);
* per sector (keys ascending) the blocks form a complete grid of row groups x column groups
  (each group a distinct qindex tuple, sorted lexicographically); the grid is one matrix, and
  its thin SVD gives that sector's singular values (descending), U rows per row group and vT
  columns per column group;
* S is diagonal on the bond aux = KET(sector keys, min(m, n) each), labels (_aux_L, _aux_R);
  U = (row legs, aux BRA) labelled (row labels, _aux_L) with rowrank = the input rowrank;
  vT = (aux KET, column legs) labelled (_aux_R, column labels) with rowrank 1.  U's blocks are
  ordered sector by sector, row groups in order; vT's likewise by column groups.

Svd_truncate (Svd_truncate_Block_UTs_internal, src/linalg/Svd_truncate.cpp:153-334): all
singular values sorted ascending; the cut value Smin is the keepdim-th largest, raised while it
is below err and more than max(mindim, 1) values remain; every value >= Smin is kept (so a
degeneracy at the cut is kept whole), sectors left empty drop out of the aux bond, and
return_err = 1 appends the largest dropped value, > 1 every dropped value (descending).

Dense UniTensors reshape to one matrix.  Every SVD is torch.linalg.svd on the tensor's GPU, in
the tensor's dtype (FP64 / complex128 for the DMRG path); singular values are real.
"""
from __future__ import annotations

from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from .device import CPU, TORCH_OF, Type_C128, Type_C64, Type_F32, Type_F64, dtype_id
from .tensor import Tensor
from .unitensor import BD_BRA, BD_KET, Bond, UniTensor, _block_layout

_REAL_OF = {Type_C128: Type_F64, Type_C64: Type_F32, Type_F64: Type_F64, Type_F32: Type_F32}


def _check_input(T: UniTensor, who: str) -> None:
    if T.rank() <= 1:
        raise ValueError("[ERROR][%s] Input UniTensor should have rank>1, but rank is %d"
                         % (who, T.rank()))
    if T.rowrank() < 1:
        raise ValueError("[ERROR][%s] Input UniTensor should have rowrank>0, but rowrank is %d"
                         % (who, T.rowrank()))
    if T.rowrank() >= T.rank():
        raise ValueError("[ERROR][%s] Input UniTensor should have rowrank<rank, but rowrank is "
                         "%d and rank is %d" % (who, T.rowrank(), T.rank()))
    if T.is_diag():
        raise ValueError("[ERROR][%s] Input UniTensor is diagonal, so the SVD is trivial and not "
                         "supported." % who)
    if T.device() == CPU:
        raise NotImplementedError("cytnx_amd implements the Device.cuda branch; move the "
                                  "UniTensor to the GPU")


def _from_blocks(bonds: Sequence[Bond], labels: Sequence[str], rowrank: int,
                 itoi: np.ndarray, blocks: List[torch.Tensor], dtype: int,
                 device: int) -> UniTensor:
    """Block UniTensor with the given block order and data."""
    out = UniTensor(bonds, labels, rowrank, dtype, device, _alloc=False)
    out._itoi = np.ascontiguousarray(itoi, dtype=np.uint64).reshape(-1, len(bonds))
    out._shapes, out._offsets, total = _block_layout(out._bonds, out._itoi)
    if blocks:
        out._storage = torch.cat([b.reshape(-1) for b in blocks]).contiguous()
    else:
        out._storage = torch.zeros(0, dtype=TORCH_OF[dtype], device=out._storage.device)
    if out._storage.numel() != total:
        raise RuntimeError("internal error: block sizes do not match the layout")
    return out


def _diag_from_blocks(bond: Bond, blocks: List[torch.Tensor], dtype: int,
                      device: int) -> UniTensor:
    S = UniTensor([bond, bond.redirect()], ["_aux_L", "_aux_R"], 1, dtype, device,
                  is_diag=True, _alloc=False)
    S._storage = torch.cat([b.reshape(-1) for b in blocks]).contiguous() if blocks else \
        torch.zeros(0, dtype=TORCH_OF[dtype], device=S._storage.device)
    return S


class _Sector:
    __slots__ = ("key", "rows", "cols", "rsz", "csz", "s", "u", "vh", "m")


def _sector_matrices(T: UniTensor) -> List[_Sector]:
    """Every sector's grid of blocks as one device matrix (sec.m), keys ascending."""
    r, rank = T.rowrank(), T.rank()
    bonds = T._bonds
    itoi = T._itoi.astype(np.int64)
    degs = [np.asarray(b.getDegeneracies(), dtype=np.int64) for b in bonds]
    keys = np.zeros(T.Nblocks(), dtype=np.int64)
    for i in range(r):
        sgn = 1 if bonds[i].type() == BD_KET else -1
        keys += sgn * np.asarray(bonds[i].qnums(), dtype=np.int64)[itoi[:, i]]
    sectors = []
    for key in sorted(set(keys.tolist())):
        blist = np.nonzero(keys == key)[0]
        rows = sorted(set(tuple(itoi[b, :r]) for b in blist))
        cols = sorted(set(tuple(itoi[b, r:]) for b in blist))
        if len(rows) * len(cols) != len(blist):
            raise RuntimeError("[Internal ERROR] the blocks of a sector do not form a grid")
        rpos = {t: i for i, t in enumerate(rows)}
        cpos = {t: i for i, t in enumerate(cols)}
        rsz = [int(np.prod([degs[i][t[i]] for i in range(r)])) for t in rows]
        csz = [int(np.prod([degs[r + j][t[j]] for j in range(rank - r)])) for t in cols]
        roff = np.concatenate([[0], np.cumsum(rsz)]).astype(int)
        coff = np.concatenate([[0], np.cumsum(csz)]).astype(int)
        M = torch.zeros((int(roff[-1]), int(coff[-1])), dtype=T._storage.dtype,
                        device=T._storage.device)
        for b in blist:
            i = rpos[tuple(itoi[b, :r])]
            j = cpos[tuple(itoi[b, r:])]
            M[roff[i]:roff[i + 1], coff[j]:coff[j + 1]] = \
                T.get_block_(int(b)).reshape(rsz[i], csz[j])
        sec = _Sector()
        sec.key, sec.rows, sec.cols, sec.rsz, sec.csz = int(key), rows, cols, rsz, csz
        sec.m = M
        sec.s = sec.u = sec.vh = None
        sectors.append(sec)
    return sectors


def _block_gesvd(T: UniTensor, is_U: bool, is_vT: bool) -> List[_Sector]:
    """Thin SVD of every sector matrix (LAPACK-style gesvd through torch.linalg.svd)."""
    sectors = _sector_matrices(T)
    for sec in sectors:
        if is_U or is_vT:
            U, s, Vh = torch.linalg.svd(sec.m, full_matrices=False)
            sec.u = U if is_U else None
            sec.vh = Vh if is_vT else None
        else:
            s = torch.linalg.svdvals(sec.m)
        sec.s = s
        sec.m = None
    return sectors


def _svd_factor(sec: _Sector, side: str) -> None:
    """The factors of _gram_factor from a thin SVD of one sector matrix."""
    U, s, Vh = torch.linalg.svd(sec.m, full_matrices=False)
    sec.s = s
    if side == "left":
        sec.u = U
        sec.vh = s.to(Vh.dtype)[:, None] * Vh
    else:
        sec.vh = Vh
        sec.u = U * s.to(U.dtype)[None, :]
    sec.m = None


def _pad_size(n: int) -> int:
    p = 8
    while p < n:
        p *= 2
    return p


def _lq_factor(sectors: List[_Sector]) -> None:
    """M = X Q per sector with Q (k x n) row-orthonormal, k = min(m, n): a QR of M^H
    (M^H = Q1 R1, X = R1^H, Q = Q1^H), sectors batched by size rounded up to powers of two
    (the batched QR's cost grows with the padding, unlike its per-call cost; measured: one call
    padded to the largest sector is 2.7x slower at chi=2048); zero padding at the end of rows
    and columns leaves the first k columns of Q1 in the true rows.  sec.s is set to ones (no
    singular values)."""
    buckets: Dict[Tuple[int, int], List[int]] = {}
    for i, sec in enumerate(sectors):
        m, n = sec.m.shape
        buckets.setdefault((_pad_size(n), _pad_size(m)), []).append(i)
    for (np_, mp), idx in sorted(buckets.items()):
        M0 = sectors[idx[0]].m
        Hb = torch.zeros((len(idx), np_, mp), dtype=M0.dtype, device=M0.device)
        for b, i in enumerate(idx):
            m, n = sectors[i].m.shape
            Hb[b, :n, :m] = sectors[i].m.conj().transpose(0, 1)
        Q1, R1 = torch.linalg.qr(Hb, mode="reduced")
        for b, i in enumerate(idx):
            sec = sectors[i]
            m, n = sec.m.shape
            k = min(m, n)
            sec.vh = Q1[b, :n, :k].conj().transpose(0, 1).contiguous()
            sec.u = R1[b, :k, :m].conj().transpose(0, 1).contiguous()
            sec.s = torch.ones(k, dtype=sec.m.real.dtype, device=sec.m.device)
            sec.m = None


def orthogonal_split(T: UniTensor) -> List[UniTensor]:
    """T = X . Q with Q right-orthonormal, no truncation: [X (rows, _aux_R), Q (_aux_R,
    columns)] with the aux bond of Gesvd (sector keys ascending, min(m, n) each) -- the
    right-canonicalisation step, by a batched LQ instead of an SVD."""
    _check_input(T, "orthogonal_split")
    if not T.is_blockform():
        raise NotImplementedError("orthogonal_split runs on block UniTensors")
    sectors = _sector_matrices(T)
    _lq_factor(sectors)
    _, X, Q = _assemble_block(T, sectors, [int(sec.s.numel()) for sec in sectors], True, True)
    X._labels[-1] = "_aux_R"
    return [X, Q]


# sectors whose Gram dimension is at most this go into a batch of their own when the largest
# Gram matrix is at least twice as large (see _gram_factor)
SMALL_GRAM = 160


def _gram_factor(sectors: List[_Sector], side: str) -> None:
    """Density-matrix factorisation of every sector matrix M (m x n), k = min(m, n), in at
    most two batched eigensolver calls: the sectors whose Gram dimension is at most SMALL_GRAM
    by themselves when the largest Gram matrix is at least twice that size, the rest together
    (a call's time grows with its padded size and with its batch count: at the centre of a
    chi=2048 chain one call over 29 sectors padded to 671 takes 27 ms, 8 sectors padded to 671
    18 ms and 21 sectors padded to 135 about 2.5 ms; splitting further adds a whole column
    sweep per call).  See _gram_batch for one call."""
    if not sectors:
        return
    left = side == "left"
    gdim = [sec.m.shape[0] if left else sec.m.shape[1] for sec in sectors]
    small = [sec for sec, g in zip(sectors, gdim) if g <= SMALL_GRAM]
    if small and len(small) < len(sectors) and max(gdim) >= 2 * SMALL_GRAM:
        _gram_batch(small, side)
        _gram_batch([sec for sec, g in zip(sectors, gdim) if g > SMALL_GRAM], side)
    else:
        _gram_batch(sectors, side)


def _gram_batch(sectors: List[_Sector], side: str) -> None:
    """Density-matrix factorisation of every sector matrix M (m x n), k = min(m, n):

    side "left":  eigh of M M^H -> U (m x k, orthonormal), sigma = sqrt(lambda), and the exact
                  product U^H M (k x n) in sec.vh;
    side "right": eigh of M^H M -> V (n x k, orthonormal), vT = V^H in sec.vh, and the exact
                  product M V (m x k) in sec.u.

    The sectors go into one batch padded to their largest Gram dimension (one eigensolver
    call: the solver sweeps the columns of the padded size once for the whole batch, so one
    padded call beats one per size class); the padding part of each Gram matrix gets distinct
    values below -1 on its
    diagonal, so its eigenvectors sort below every true one (the true eigenvalues are >= 0) and
    the top k eigenvectors live in the true rows.  Eigenpairs come out in descending order.
    sigma is accurate to about eps * sigma_max^2 / sigma (absolute); the factor kept orthonormal
    is orthonormal to working precision and the other factor is a product, never a division by
    sigma."""
    if not sectors:
        return
    left = side == "left"
    dims = [sec.m.shape for sec in sectors]
    gs = [d[0] if left else d[1] for d in dims]
    os_ = [d[1] if left else d[0] for d in dims]
    gp, op = max(gs), max(os_)
    M0 = sectors[0].m
    Mb = torch.zeros((len(sectors), gp, op) if left else (len(sectors), op, gp),
                     dtype=M0.dtype, device=M0.device)
    for b, sec in enumerate(sectors):
        m, n = sec.m.shape
        Mb[b, :m, :n] = sec.m
    MbH = Mb.conj().transpose(1, 2).contiguous()
    G = torch.bmm(Mb, MbH) if left else torch.bmm(MbH, Mb)
    # distinct negative values on the padding diagonal: no repeated eigenvalues from padding
    gidx = torch.arange(gp, device=G.device)
    glen = torch.tensor(gs, device=G.device)[:, None]
    padv = -1.0 - (gidx[None, :] - glen).clamp(min=0).to(G.real.dtype) / gp
    diag = torch.diagonal(G, dim1=1, dim2=2)
    diag.copy_(torch.where(gidx[None, :] < glen, diag, padv.to(G.dtype)))
    try:
        lam, vec = torch.linalg.eigh(G)
    except torch.linalg.LinAlgError:
        # the batched eigensolver did not converge on some Gram matrix (many repeated
        # eigenvalues, e.g. a rank-deficient sector): thin SVD of every sector instead
        for sec in sectors:
            _svd_factor(sec, side)
        return
    kmax = max(min(d) for d in dims)
    sig = torch.sqrt(torch.clamp(lam.flip(-1)[:, :kmax], min=0.0))   # descending
    vd = vec.flip(-1)[:, :, :kmax]                                     # matching vectors
    if left:
        other = torch.bmm(vd.conj().transpose(1, 2), Mb)               # U^H M  (B, kmax, op)
    else:
        other = torch.bmm(Mb, vd)                                      # M V    (B, op, kmax)
    for b, sec in enumerate(sectors):
        m, n = sec.m.shape
        k = min(m, n)
        sec.s = sig[b, :k]
        if left:
            sec.u = vd[b, :m, :k]
            sec.vh = other[b, :k, :n]
        else:
            sec.vh = torch.conj_physical(vd[b, :n, :k].transpose(0, 1))
            sec.u = other[b, :m, :k]
        sec.m = None


def truncated_split(T: UniTensor, keepdim: int, absorb: str, err: float = 0.0,
                    mindim: int = 1) -> List[UniTensor]:
    """The truncated factorisation T ~ X . Y of DMRG's bond update, with the block cut rule of
    Svd_truncate on the singular values, by the density-matrix form (``_gram_factor``):

    absorb "right": X = U (rows, _aux_L; left-orthonormal), Y = S vT (_aux_L, columns), as
                    Svd_truncate followed by Contract(s, vT);
    absorb "left":  X = U S (rows, _aux_R), Y = vT (_aux_R, columns; right-orthonormal), as
                    Contract(U, s) and vT.

    Returns [S, X, Y] (S the kept singular values, diagonal)."""
    _check_input(T, "truncated_split")
    if not T.is_blockform():
        raise NotImplementedError("truncated_split runs on block UniTensors")
    if absorb not in ("right", "left"):
        raise ValueError("absorb is 'right' or 'left'")
    if keepdim < 1:
        raise ValueError("keepdim must be >=1")
    sectors = _sector_matrices(T)
    _gram_factor(sectors, "left" if absorb == "right" else "right")
    lens = [int(sec.s.numel()) for sec in sectors]
    svals = torch.cat([sec.s for sec in sectors]).detach().cpu().numpy()   # one transfer
    Smin, _, _ = _cut(svals, keepdim, err, mindim)
    bounds = np.concatenate([[0], np.cumsum(lens)]).astype(int)
    keep = [int(np.count_nonzero(svals[bounds[i]:bounds[i + 1]] >= Smin))
            for i in range(len(sectors))]
    S, X, Y = _assemble_block(T, sectors, keep, True, True)
    if absorb == "right":
        Y._labels[0] = "_aux_L"
    else:
        X._labels[-1] = "_aux_R"
    return [S, X, Y]


def _assemble_block(T: UniTensor, sectors: List[_Sector], keep: List[int], is_U: bool,
                    is_vT: bool) -> List[UniTensor]:
    """S, U, vT from the sectors, keeping the first keep[i] singular values of sector i and
    dropping sectors with keep[i] == 0."""
    r, rank = T.rowrank(), T.rank()
    dt, dev = T.dtype(), T.device()
    live = [i for i, k in enumerate(keep) if k > 0]
    aux = Bond(BD_KET, [[sectors[i].key] for i in live], [keep[i] for i in live])
    S = _diag_from_blocks(aux, [sectors[i].s[:keep[i]] for i in live], _REAL_OF[dt], dev)
    out = [S]
    if is_U:
        itoi, blocks = [], []
        for new_i, i in enumerate(live):
            sec = sectors[i]
            off = 0
            for t, m in zip(sec.rows, sec.rsz):
                shp = [int(T._bonds[a].getDegeneracies()[t[a]]) for a in range(r)]
                blocks.append(sec.u[off:off + m, :keep[i]].reshape(shp + [keep[i]]))
                itoi.append(list(t) + [new_i])
                off += m
        out.append(_from_blocks(list(T._bonds[:r]) + [aux.redirect()],
                                list(T._labels[:r]) + ["_aux_L"], r,
                                np.asarray(itoi, dtype=np.uint64).reshape(-1, r + 1), blocks,
                                dt, dev))
    if is_vT:
        itoi, blocks = [], []
        for new_i, i in enumerate(live):
            sec = sectors[i]
            off = 0
            for t, n in zip(sec.cols, sec.csz):
                shp = [int(T._bonds[r + a].getDegeneracies()[t[a]]) for a in range(rank - r)]
                blocks.append(sec.vh[:keep[i], off:off + n].reshape([keep[i]] + shp))
                itoi.append([new_i] + list(t))
                off += n
        out.append(_from_blocks([aux] + list(T._bonds[r:]), ["_aux_R"] + list(T._labels[r:]),
                                1, np.asarray(itoi, dtype=np.uint64).reshape(-1, rank - r + 1),
                                blocks, dt, dev))
    return out


def _dense_matrix(T: UniTensor) -> Tuple[torch.Tensor, List[int]]:
    t = T._block.contiguous()
    shp = t.shape()
    m = int(np.prod(shp[:T.rowrank()]))
    return t.torch().reshape(m, -1), shp


def _assemble_dense(T: UniTensor, U, s, Vh, k: int, is_U: bool, is_vT: bool) -> List[UniTensor]:
    r = T.rowrank()
    shp = T._block.shape()
    dev = T.device()
    rdt = _REAL_OF[T.dtype()]
    S = UniTensor([Bond(k), Bond(k)], ["_aux_L", "_aux_R"], 1, rdt, dev, is_diag=True)
    S._block = Tensor.from_storage(s[:k].contiguous(), [k])
    out = [S]
    if is_U:
        Ut = UniTensor([Bond(d) for d in shp[:r]] + [Bond(k)],
                       list(T._labels[:r]) + ["_aux_L"], r, T.dtype(), dev)
        Ut._block = Tensor.from_storage(U[:, :k].contiguous(), list(shp[:r]) + [k])
        out.append(Ut)
    if is_vT:
        vT = UniTensor([Bond(k)] + [Bond(d) for d in shp[r:]],
                       ["_aux_R"] + list(T._labels[r:]), 1, T.dtype(), dev)
        vT._block = Tensor.from_storage(Vh[:k, :].contiguous(), [k] + list(shp[r:]))
        out.append(vT)
    return out


def Gesvd(T: UniTensor, is_U: bool = True, is_vT: bool = True) -> List[UniTensor]:
    """[S, U (if is_U), vT (if is_vT)] (linalg::Gesvd, src/linalg/Gesvd.cpp)."""
    _check_input(T, "Gesvd")
    if T.is_blockform():
        sectors = _block_gesvd(T, is_U, is_vT)
        return _assemble_block(T, sectors, [int(sec.s.numel()) for sec in sectors], is_U, is_vT)
    M, _ = _dense_matrix(T)
    if is_U or is_vT:
        U, s, Vh = torch.linalg.svd(M, full_matrices=False)
    else:
        U, s, Vh = None, torch.linalg.svdvals(M), None
    return _assemble_dense(T, U, s, Vh, int(s.numel()), is_U, is_vT)


def Svd(T: UniTensor, is_UvT: bool = True) -> List[UniTensor]:
    """[S, U, vT] or [S] (linalg::Svd)."""
    return Gesvd(T, is_UvT, is_UvT)


def _cut(svals: np.ndarray, keepdim: int, err: float, mindim: int
         ) -> Tuple[float, int, np.ndarray]:
    """(Smin, number dropped, all values ascending) of the block truncation rule."""
    Sall = np.sort(svals)  # ascending
    n = Sall.size
    keep = keepdim
    if keep < n:
        smidx = n - keep
    else:
        keep, smidx = n, 0
    Smin = Sall[smidx]
    lo = max(1, mindim)
    while Smin < err and keep > lo:
        keep -= 1
        smidx += 1
        Smin = Sall[smidx]
    while smidx > 0 and not (Sall[smidx - 1] < Smin):
        smidx -= 1
    return float(Smin), smidx, Sall


def _cut_dense(s_desc: np.ndarray, keepdim: int, err: float, mindim: int) -> int:
    """Kept count of the dense rule (memcpyTruncation, src/backend/linalg_internal_cpu/
    memcpyTruncation.cpp:13-60): the first min(n, keepdim) values, then trailing values below
    err dropped while more than max(mindim, 1) remain; no degeneracy rule."""
    mindim = max(1, mindim)
    k = min(int(s_desc.size), int(keepdim))
    for i in range(k - 1, -1, -1):
        if s_desc[i] < err and k - 1 >= mindim:
            k -= 1
        else:
            break
    return max(k, 1) if s_desc.size else 0


def _err_tensor(Sall: np.ndarray, dropped: int, return_err: int, dtype: int,
                device: int) -> UniTensor:
    if return_err == 1:
        T = UniTensor([], [], 0, dtype, device)
        v = Sall[dropped - 1] if dropped else 0.0
        T._block = Tensor.from_storage(torch.tensor([v], dtype=TORCH_OF[dtype],
                                                    device=T._block.storage().device), [])
        return T
    vals = Sall[:dropped][::-1].copy()
    T = UniTensor([Bond(dropped)], ["0"], 0, dtype, device)
    T._block = Tensor.from_storage(torch.tensor(vals, dtype=TORCH_OF[dtype],
                                                device=T._block.storage().device), [dropped])
    return T


def _cut_blockdim(blocks: Sequence[np.ndarray], keepdim: int, min_blockdim: Sequence[int],
                  err: float, mindim: int) -> Tuple[List[int], np.ndarray, int]:
    """Kept count per block, the candidate values (ascending) and how many of them are dropped,
    for the block rule with a per-block floor (Svd_truncate_Block_UTs_internal with
    min_blockdim, src/linalg/Svd_truncate.cpp:357-520; the same steps in
    src/linalg/Gesvd_truncate.cpp:431-590):

    1. block b keeps its min_blockdim[b] largest values outright (all of them when it has no
       more), and keepdim and mindim are reduced by what it keeps so; its other values are
       candidates;
    2. over the candidates the rule of Svd_truncate runs with the reduced keepdim / mindim (the
       floor is max(mindim, 1) if no block has a floor, else max(mindim, 0); the err loop may
       end at zero kept candidates), and a block also keeps every candidate >= the cut value;
    3. with no candidates at all every block is kept whole, and with the reduced keepdim <= 0
       every candidate is dropped.
    min_blockdim with one element applies to every block."""
    nb = len(blocks)
    mbd = [int(x) for x in min_blockdim]
    if not mbd:
        raise ValueError("min_blockdim must not be empty; use the overload without min_blockdim "
                         "if no per-block floor is needed.")
    if len(mbd) == 1:
        mbd = mbd * nb
    if len(mbd) != nb:
        raise ValueError("min_blockdim must have the same number of elements as blocks in the "
                         "singular value UniTensor")
    keep_dim, min_dim = int(keepdim), max(int(mindim), 1)
    cand, any_floor = [], False
    for s, m in zip(blocks, mbd):
        if m < 1:
            cand.append(s)
            continue
        any_floor = True
        if s.size <= m:
            keep_dim -= s.size
            min_dim -= s.size
            continue
        cand.append(s[m:])
        keep_dim -= m
        min_dim -= m
    if not cand:
        return [int(s.size) for s in blocks], np.zeros(0), 0
    Sall = np.sort(np.concatenate(cand))
    n = Sall.size
    Smin = 0.0
    if keep_dim > 0:
        min_dim = max(min_dim, 0 if any_floor else 1)
        if keep_dim < n:
            smidx = n - keep_dim
        else:
            keep_dim, smidx = n, 0
        Smin = Sall[smidx]
        while Smin < err and keep_dim > min_dim:
            keep_dim -= 1
            if keep_dim == 0:
                break
            smidx += 1
            Smin = Sall[smidx]
        if keep_dim == 0:
            dropped = n
        else:
            while smidx > 0 and not (Sall[smidx - 1] < Smin):
                smidx -= 1
            dropped = smidx
    else:
        dropped = n
    keep = []
    for s, m in zip(blocks, mbd):
        k = min(m, int(s.size))
        if keep_dim > 0:
            for i in range(int(s.size), m, -1):
                if s[i - 1] >= Smin:
                    k = i
                    break
        keep.append(max(k, 0))
    return keep, Sall, dropped


def _truncate(T: UniTensor, keepdim: int, min_blockdim, err: float, is_U: bool, is_vT: bool,
              return_err: int, mindim: int, who: str) -> List[UniTensor]:
    """The block / dense truncated SVD shared by Svd_truncate and Gesvd_truncate."""
    _check_input(T, who)
    if keepdim < 1:
        raise ValueError("[ERROR][%s] keepdim must be >=1" % who)
    if return_err < 0:
        raise ValueError("[ERROR][%s] return_err cannot be negative" % who)
    rdt = _REAL_OF[T.dtype()]
    if T.is_blockform():
        sectors = _block_gesvd(T, is_U, is_vT)
        svals = [sec.s.detach().cpu().numpy() for sec in sectors]
        if min_blockdim is None:
            Smin, dropped, Sall = _cut(np.concatenate(svals) if svals else np.zeros(0),
                                       keepdim, err, mindim)
            keep = [int(np.count_nonzero(s >= Smin)) for s in svals]
        else:
            keep, Sall, dropped = _cut_blockdim(svals, keepdim, min_blockdim, err, mindim)
        out = _assemble_block(T, sectors, keep, is_U, is_vT)
    else:
        if min_blockdim is not None:
            if len(min_blockdim) != 1:
                raise ValueError("[ERROR][%s] min_blockdim must have one element for dense "
                                 "UniTensor" % who)
            mindim = max(int(mindim), int(min_blockdim[0]))
        M, _ = _dense_matrix(T)
        if is_U or is_vT:
            U, s, Vh = torch.linalg.svd(M, full_matrices=False)
        else:
            U, s, Vh = None, torch.linalg.svdvals(M), None
        sh = s.detach().cpu().numpy()
        k = _cut_dense(sh, keepdim, err, mindim)
        dropped, Sall = sh.size - k, np.sort(sh)
        out = _assemble_dense(T, U, s, Vh, k, is_U, is_vT)
    if return_err:
        out.append(_err_tensor(Sall, dropped, int(return_err), rdt, T.device()))
    return out


def _split_overload(args: tuple, kwargs: dict) -> tuple:
    """(min_blockdim or None, remaining positional arguments): the overloads with a per-block
    floor take a sequence of integers as their third argument."""
    if "min_blockdim" in kwargs:
        return list(kwargs.pop("min_blockdim")), args
    if args and isinstance(args[0], (list, tuple, np.ndarray)):
        return list(args[0]), args[1:]
    return None, args


def Svd_truncate(T: UniTensor, keepdim: int, *args, **kwargs) -> List[UniTensor]:
    """[S, U, vT] (+ [err]) keeping at most keepdim singular values (see module docstring).

    Svd_truncate(T, keepdim, err=0, is_UvT=True, return_err=0, mindim=1) or, with a per-block
    floor, Svd_truncate(T, keepdim, min_blockdim, err=0, is_UvT=True, return_err=0, mindim=1)
    (include/linalg.hpp:753-810)."""
    mbd, args = _split_overload(args, kwargs)
    names = ("err", "is_UvT", "return_err", "mindim")
    vals = dict(zip(names, (0.0, True, 0, 1)))
    vals.update(zip(names, args))
    vals.update(kwargs)
    return _truncate(T, keepdim, mbd, float(vals["err"]), bool(vals["is_UvT"]),
                     bool(vals["is_UvT"]), int(vals["return_err"]), int(vals["mindim"]),
                     "Svd_truncate")


def Gesvd_truncate(T: UniTensor, keepdim: int, *args, **kwargs) -> List[UniTensor]:
    """[S, U (if is_U), vT (if is_vT)] (+ [err]): Svd_truncate with separate U / vT flags
    (linalg::Gesvd_truncate, include/linalg.hpp:824-845; src/linalg/Gesvd_truncate.cpp:222-430:
    the same cut rule on the singular values of Gesvd).

    Gesvd_truncate(T, keepdim, err=0, is_U=True, is_vT=True, return_err=0, mindim=1) or
    Gesvd_truncate(T, keepdim, min_blockdim, err=0, is_U=True, is_vT=True, return_err=0,
    mindim=1)."""
    mbd, args = _split_overload(args, kwargs)
    names = ("err", "is_U", "is_vT", "return_err", "mindim")
    vals = dict(zip(names, (0.0, True, True, 0, 1)))
    vals.update(zip(names, args))
    vals.update(kwargs)
    return _truncate(T, keepdim, mbd, float(vals["err"]), bool(vals["is_U"]),
                     bool(vals["is_vT"]), int(vals["return_err"]), int(vals["mindim"]),
                     "Gesvd_truncate")


__all__ = ["Gesvd", "Svd", "Svd_truncate", "Gesvd_truncate"]
