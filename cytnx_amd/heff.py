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
((L.A).W).R; its three steps are strided-batch GEMM plans that read and write every tensor in
place.

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
                        _Plan, contract_meta, meta_elems, plan_contraction)


class _Chain:
    def __init__(self, steps, device: int):
        """steps: list of (plan, fixed operand (a UniTensor or anything with a flat
        ``_storage``), fixed_is_left, output meta UniTensor or output element count)."""
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
            _lib.i64_array([s[3] if isinstance(s[3], int) else meta_elems(s[3])
                            for s in steps]), ctypes.byref(h)))
        self.handle = h.value
        f = ctypes.c_double(0)
        _lib.check(_lib.lib().cxa_chain_flops(self.handle, ctypes.byref(f)))
        self.flops = f.value
        self.prec = None

    def layout(self, step: int) -> int:
        """Operand layouts of a step (cxa_chain_layout): bit 0 packed fixed B, bit 1 packed
        running A, bit 2 outputs written packed for the next step, bit 3 (step 0, FP32 prepared)
        FP32 ground-state Lanczos runs feed it a packed FP32 copy of the Lanczos vector."""
        f = ctypes.c_int(0)
        _lib.check(_lib.lib().cxa_chain_layout(self.handle, int(step), ctypes.byref(f)))
        return f.value

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


MIX_SLICES = 8  # CXA_MIX_SLICES


class MixUnsupported(ValueError):
    """The operator does not fit a mix launch (an item with more than MIX_SLICES slices)."""


def _row_strides(shape) -> np.ndarray:
    st = np.ones(len(shape), dtype=np.int64)
    for d in range(len(shape) - 2, -1, -1):
        st[d] = st[d + 1] * int(shape[d + 1])
    return st


def _sector_offsets(b: Bond) -> np.ndarray:
    degs = b.getDegeneracies()
    return np.concatenate([[0], np.cumsum(degs)[:-1]]).astype(np.int64)


def psi_is_left(Y: UniTensor, labY: dict) -> bool:
    """Whether a mix launch writing Y takes psi as its left operand: Y's l' leg comes after its
    r leg (l' contiguous), so each item computes the transposed tile."""
    return Y.labels().index(labY["lp"]) > Y.labels().index(labY["r"])


def mix_plan_arrays(Lp: UniTensor, labL: dict, psi: UniTensor, labP: dict, W12: np.ndarray,
                    Y: UniTensor, labY: dict, live: Optional[np.ndarray] = None,
                    read: Optional[np.ndarray] = None):
    """Task arrays of the mix launch Y = W12 applied to (Lp . psi), never storing Lp . psi.

    Lp: (l, w, l') environment blocks; psi: the two-site state (l, s1, s2, r); W12: the dense
    two-site MPO with axes (w, s1, s1', w'', s2, s2'); Y: the output layout, blocks with the
    legs (l', s1', s2', w'', r) in any order, r or l' contiguous.  labL / labP / labY map the
    roles (l, w, lp / l, s1, s2, r / lp, s1p, s2p, w2, r) to each tensor's labels.  One item
    per (l', r) sector pair of Y; its slices are the products
    X[l', r] = sum_l Lp[l, w, l'] psi[l, s1, s2, r]  for every (w, s1, s2) element that
    reaches one of the item's outputs, and each output Y[l', s1', s2', w'', r] is the sum over
    slices of W12[w, s1, s1', w'', s2, s2'] * X.

    The contiguous axis of Y is the launch tile's column axis.  With r contiguous (r after l'
    in Y's label order) the items compute X[l', r] with A = Lp and B = psi; with l' contiguous
    (l' after r) they compute the transposed tile X[r, l'] with A = psi and B = Lp, and the
    launch takes psi as its left operand.

    Returns (items (n,4), slices (m,7), outs (k,2), coef (k,MIX_SLICES)) as int64 / float64
    arrays, or raises MixUnsupported.  With live (a bool array of Y.Nblocks()), marks the Y
    blocks that some slice reaches (the others are zero for every input).  With read (a bool
    array of Y.Nblocks(): the blocks the next launch reads), the plan is made in one pass for
    the pruned layout subset_blocks(Y, keep), keep = reached and read: the other blocks get no
    outputs, the offsets address the pruned layout, and live (if given) receives keep."""
    pL = {k: Lp.labels().index(labL[k]) for k in ("l", "w", "lp")}
    pP = {k: psi.labels().index(labP[k]) for k in ("l", "s1", "s2", "r")}
    pY = {k: Y.labels().index(labY[k]) for k in ("lp", "s1p", "s2p", "w2", "r")}
    psi_left = psi_is_left(Y, labY)
    col = "lp" if psi_left else "r"
    offw = _sector_offsets(Lp.bonds()[pL["w"]])
    offs1 = _sector_offsets(psi.bonds()[pP["s1"]])
    offs2 = _sector_offsets(psi.bonds()[pP["s2"]])
    offs1p = _sector_offsets(Y.bonds()[pY["s1p"]])
    offs2p = _sector_offsets(Y.bonds()[pY["s2p"]])
    offw2 = _sector_offsets(Y.bonds()[pY["w2"]])
    stL = [_row_strides(s) for s in Lp._shapes]
    stP = [_row_strides(s) for s in psi._shapes]
    stY = [_row_strides(s) for s in Y._shapes]
    L_by = {}
    for bi, t in enumerate(Lp._itoi):
        L_by.setdefault((int(t[pL["l"]]), int(t[pL["lp"]])), []).append(bi)
    P_by = {}
    for bi, t in enumerate(psi._itoi):
        P_by.setdefault(int(t[pP["r"]]), []).append(bi)
    Y_by = {}
    for yi, t in enumerate(Y._itoi):
        Y_by.setdefault((int(t[pY["lp"]]), int(t[pY["r"]])), []).append(yi)
    items, slices, outs, coefs, out_block, records = [], [], [], [], [], []
    keep_blocks = np.zeros(Y.Nblocks(), dtype=bool)
    for (qap, qc), ylist in Y_by.items():
        cand = []
        for pi in P_by.get(qc, []):
            tp, shp = psi._itoi[pi], psi._shapes[pi]
            stp = stP[pi]
            qa = int(tp[pP["l"]])
            for li in L_by.get((qa, qap), []):
                tl, shl = Lp._itoi[li], Lp._shapes[li]
                stl = stL[li]
                K = int(shl[pL["l"]])
                if K != int(shp[pP["l"]]):
                    raise ValueError("Heff2 mix plan: L and psi disagree on a sector of l")
                for iw in range(int(shl[pL["w"]])):
                    for i1 in range(int(shp[pP["s1"]])):
                        for i2 in range(int(shp[pP["s2"]])):
                            a_l = int(Lp._offsets[li] + iw * stl[pL["w"]])
                            a_p = int(psi._offsets[pi] + i1 * stp[pP["s1"]]
                                      + i2 * stp[pP["s2"]])
                            if psi_left:  # X[r, l'] = psi^T . Lp
                                desc = (a_p, a_l, int(stp[pP["r"]]), int(stp[pP["l"]]),
                                        int(stl[pL["l"]]), int(stl[pL["lp"]]), K)
                                mn = (int(shp[pP["r"]]), int(shl[pL["lp"]]))
                            else:  # X[l', r] = Lp^T . psi
                                desc = (a_l, a_p, int(stl[pL["lp"]]), int(stl[pL["l"]]),
                                        int(stp[pP["l"]]), int(stp[pP["r"]]), K)
                                mn = (int(shl[pL["lp"]]), int(shp[pP["r"]]))
                            g = (int(offw[int(tl[pL["w"]])] + iw),
                                 int(offs1[int(tp[pP["s1"]])] + i1),
                                 int(offs2[int(tp[pP["s2"]])] + i2))
                            cand.append((desc, g, mn[0], mn[1]))
        olist = []
        oblock = []
        M = N = None
        for yi in ylist:
            ty, shy = Y._itoi[yi], Y._shapes[yi]
            sty = stY[yi]
            if sty[pY[col]] != 1:
                raise ValueError("Heff2 mix plan: %s must be the contiguous axis of Y" % col)
            row = "r" if psi_left else "lp"
            M, N = int(shy[pY[row]]), int(shy[pY[col]])
            for i1p in range(int(shy[pY["s1p"]])):
                for i2p in range(int(shy[pY["s2p"]])):
                    for j in range(int(shy[pY["w2"]])):
                        c_off = int(i1p * sty[pY["s1p"]] + i2p * sty[pY["s2p"]]
                                    + j * sty[pY["w2"]])  # within block yi
                        olist.append((c_off, int(sty[pY[row]]),
                                      (int(offs1p[int(ty[pY["s1p"]])] + i1p),
                                       int(offw2[int(ty[pY["w2"]])] + j),
                                       int(offs2p[int(ty[pY["s2p"]])] + i2p))))
                        oblock.append(yi)
        C = np.zeros((len(olist), len(cand)))
        for o, (_, _, (s1p, w2, s2p)) in enumerate(olist):
            for s, (_, (w, s1, s2), _, _) in enumerate(cand):
                C[o, s] = W12[w, s1, s1p, w2, s2, s2p]
        reached = {yi for o, yi in enumerate(oblock) if np.any(C[o] != 0.0)}
        if live is not None and read is None:
            for yi in reached:
                live[yi] = True
        if read is not None:
            # one pass for the pruned layout: only blocks that are reached and read keep their
            # outputs (all of them, zero rows included: the block is read)
            kept = [o for o, yi in enumerate(oblock) if yi in reached and read[yi]]
            for yi in reached:
                if read[yi]:
                    keep_blocks[yi] = True
            if not kept:
                continue
            olist = [olist[o] for o in kept]
            oblock = [oblock[o] for o in kept]
            C = C[kept]
        keep = np.any(C != 0.0, axis=0) if len(olist) else np.zeros(len(cand), dtype=bool)
        cand = [c for c, k in zip(cand, keep) if k]
        C = C[:, keep]
        if len(cand) > MIX_SLICES:
            raise MixUnsupported("an item of the two-site operator has %d slices (at most %d)"
                                 % (len(cand), MIX_SLICES))
        for c in cand:
            if (c[2], c[3]) != (M, N):
                raise ValueError("Heff2 mix plan: slice shape differs from its outputs'")
        rows = []
        for o in range(len(olist)):
            row = np.zeros(MIX_SLICES)
            row[:len(cand)] = C[o]
            rows.append(row)
        # items in the order of their first block in the layout the plan addresses
        records.append((min(oblock) if oblock else len(records), (M, N, len(cand), len(olist)),
                        [c[0] for c in cand], [(c_off, ldc) for c_off, ldc, _ in olist],
                        list(oblock), rows))
    if read is not None:
        records.sort(key=lambda rec: rec[0])
    for _, item, sl, ou, ob, rows in records:
        items.append(item)
        slices.extend(sl)
        outs.extend(ou)
        out_block.extend(ob)
        coefs.extend(rows)
    # block offsets of the layout the outputs address
    if read is not None:
        sizes = np.prod(Y._shapes.astype(np.int64), axis=1) if Y.Nblocks() else \
            np.zeros(0, dtype=np.int64)
        base = np.zeros(Y.Nblocks(), dtype=np.int64)
        ks = sizes * keep_blocks
        if len(ks):
            base[1:] = np.cumsum(ks)[:-1]
        if live is not None:
            live[:] = keep_blocks
    else:
        base = np.asarray(Y._offsets, dtype=np.int64)
    outs = [(int(base[b]) + c, l) for (c, l), b in zip(outs, out_block)]
    as64 = lambda rows, w: np.asarray(rows, dtype=np.int64).reshape(-1, w)
    return (as64(items, 4), as64(slices, 7), as64(outs, 2),
            np.asarray(coefs, dtype=np.float64).reshape(-1, MIX_SLICES))


def mix_plan_arrays_native(Lp: UniTensor, labL: dict, psi: UniTensor, labP: dict,
                           W12: np.ndarray, Y: UniTensor, labY: dict,
                           live: Optional[np.ndarray] = None, read: Optional[np.ndarray] = None):
    """mix_plan_arrays computed by the library's C++ planner (cxa_heff2_mix_plan): the same
    arguments and the same arrays, element for element."""
    pL = [Lp.labels().index(labL[k]) for k in ("l", "w", "lp")]
    pP = [psi.labels().index(labP[k]) for k in ("l", "s1", "s2", "r")]
    pY = [Y.labels().index(labY[k]) for k in ("lp", "s1p", "s2p", "w2", "r")]
    bonds = [Lp.bonds()[pL[1]], psi.bonds()[pP[1]], psi.bonds()[pP[2]], Y.bonds()[pY[1]],
             Y.bonds()[pY[2]], Y.bonds()[pY[3]]]
    secs = [_sector_offsets(b) for b in bonds]
    sec_off = np.ascontiguousarray(np.concatenate(secs).astype(np.int64))
    sec_n = np.asarray([len(x) for x in secs], dtype=np.int64)
    W = np.ascontiguousarray(W12, dtype=np.float64)
    wd = np.asarray(W.shape, dtype=np.int64)

    def meta(T, rank):
        it = np.ascontiguousarray(np.asarray(T._itoi, dtype=np.int64).reshape(-1, rank))
        sh = np.ascontiguousarray(np.asarray(T._shapes, dtype=np.int64).reshape(-1, rank))
        of = np.ascontiguousarray(np.asarray(T._offsets, dtype=np.int64))
        return it, sh, of

    iL, sL, oL = meta(Lp, 3)
    iP, sP, oP = meta(psi, 4)
    iY, sY, oY = meta(Y, 5)
    nbY = Y.Nblocks()
    cap_items = max(1, nbY)
    cap_slices = MIX_SLICES * cap_items
    cap_outs = max(1, int(np.sum(np.prod(sY[:, [pY[1], pY[2], pY[3]]], axis=1)))) if nbY else 1
    items = np.zeros((cap_items, 4), dtype=np.int64)
    slices = np.zeros((cap_slices, 7), dtype=np.int64)
    outs = np.zeros((cap_outs, 2), dtype=np.int64)
    coef = np.zeros((cap_outs, MIX_SLICES), dtype=np.float64)
    rd = None if read is None else np.ascontiguousarray(read, dtype=np.uint8)
    lv = np.zeros(max(1, nbY), dtype=np.uint8)
    counts = np.zeros(3, dtype=np.int64)
    ip = lambda a: a.ctypes.data_as(_lib.i64p)
    rc = _lib.lib().cxa_heff2_mix_plan(
        Lp.Nblocks(), ip(iL), ip(sL), ip(oL), _lib.int_array(pL),
        psi.Nblocks(), ip(iP), ip(sP), ip(oP), _lib.int_array(pP),
        nbY, ip(iY), ip(sY), ip(oY), _lib.int_array(pY),
        ip(sec_off), ip(sec_n), W.ctypes.data_as(_lib.dblp), ip(wd),
        rd.ctypes.data if rd is not None else None, cap_items, cap_slices, cap_outs,
        ip(items), ip(slices), ip(outs), coef.ctypes.data_as(_lib.dblp), lv.ctypes.data,
        ip(counts))
    if rc == _lib.ERR_UNSUPPORTED:
        raise MixUnsupported(_lib.lib().cxa_last_error().decode(errors="replace"))
    _lib.check(rc)
    if live is not None:
        live[:] = lv[:nbY].astype(bool)
    ni, ns, no = (int(c) for c in counts)
    return items[:ni].copy(), slices[:ns].copy(), outs[:no].copy(), coef[:no].copy()


def subset_blocks(T: UniTensor, keep: np.ndarray) -> UniTensor:
    """Metadata of T restricted to the blocks with keep[i] (same order, packed layout)."""
    out = T.clone_meta()
    out._itoi = np.ascontiguousarray(T._itoi[keep])
    out._shapes = np.ascontiguousarray(T._shapes[keep])
    sizes = np.prod(out._shapes.astype(np.int64), axis=1) if len(out._shapes) else \
        np.zeros(0, dtype=np.int64)
    out._offsets = np.zeros(len(sizes), dtype=np.int64)
    if len(sizes):
        out._offsets[1:] = np.cumsum(sizes)[:-1]
    out._storage = T._storage[:0]
    return out


def check_mix_bounds(arrays, na: int, nb: int, nc: int) -> None:
    """Every element a mix launch reads (A: na elements, B: nb) or writes (C: nc) lies inside
    its buffer; raises ValueError otherwise (checked on the host before any launch)."""
    items, slices, outs, _ = arrays
    ns = items[:, 2].astype(np.int64)
    no = items[:, 3].astype(np.int64)
    M = np.repeat(items[:, 0], ns)
    N = np.repeat(items[:, 1], ns)
    if len(slices):
        a_off, b_off, sam, sak, sbk, sbn, K = (slices[:, i] for i in range(7))
        live = (K > 0) & (M > 0) & (N > 0)
        amax = a_off + (M - 1) * sam + (K - 1) * sak
        bmax = b_off + (K - 1) * sbk + (N - 1) * sbn
        if np.any(live & ((amax >= na) | (bmax >= nb) | (a_off < 0) | (b_off < 0))):
            raise ValueError("mix plan reads outside its operands")
    if len(outs):
        Mo = np.repeat(items[:, 0], no)
        No = np.repeat(items[:, 1], no)
        cmax = outs[:, 0] + (Mo - 1) * outs[:, 1] + No - 1
        if np.any((Mo > 0) & (No > 0) & ((cmax >= nc) | (outs[:, 0] < 0))):
            raise ValueError("mix plan writes outside its output")


class _MixPlan(_Plan):
    """A cxa_mix_contract_plan handle (exec / info / destroy are the block-contract ones)."""

    def __init__(self, arrays, device: int):
        items, slices, outs, coef = (np.ascontiguousarray(a) for a in arrays)
        h = ctypes.c_void_p(0)
        torch.cuda.set_device(device)
        _lib.check(_lib.lib().cxa_mix_contract_plan(
            3, int(items.shape[0]), items.ctypes.data_as(_lib.i64p), int(slices.shape[0]),
            slices.ctypes.data_as(_lib.i64p), int(outs.shape[0]), outs.ctypes.data_as(_lib.i64p),
            coef.ctypes.data_as(_lib.dblp), ctypes.byref(h)))
        super().__init__(h.value, device)


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

    def set_tiles(self, choice: Sequence[dict]) -> None:
        """Apply a per-step tiling choice (as returned by tune) at the current precision."""
        if len(choice) != len(self.chain.steps):
            raise ValueError("tiling choice has %d steps, chain has %d"
                             % (len(choice), len(self.chain.steps)))
        for plan, c in zip(self.chain._plans, choice):
            _lib.check(_lib.lib().cxa_block_contract_set_config(
                plan.handle, int(self.precision), int(c["kind"]), int(c["split"])))

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
    W12 = M1.M2 (over the shared MPO bond).  With ``mix`` as well (default) the matvec is two
    launches: a mix launch that forms the (w, s1, s2) slices of L.psi per (l', r) sector pair in
    registers and writes their W12-weighted sums Y(l', s1', s2', w'', r) directly, then the
    grouped GEMM Y.R over (w'', r) -- neither L.psi nor the M1 intermediate is stored.  Without
    ``mix`` (or when an item would exceed the mix launch's slice bound) the matvec is three
    grouped launches L.psi, .W12, .R.  ``flops`` is always the algorithmic count of the
    reference's four-step order L.psi -> M1 -> M2 -> R."""

    def __init__(self, psi: UniTensor, L: UniTensor, M1: UniTensor, M2: UniTensor,
                 R: UniTensor, fuse_mpo: bool = True, mix: bool = True):
        for T in (psi, L, M1, M2, R):
            if not T.is_blockform():
                raise TypeError("Heff2 takes U(1) block tensors")
            if T.dtype() not in (1, 3):
                raise TypeError("Heff2 runs FP64 or complex128 tensors")
        # complex128 operands (a complex state or environments): every tensor complex, the
        # three-launch chain in FP64 (the mix launch and the FP32 modes are real-only)
        dt = 1 if any(T.dtype() == 1 for T in (psi, L, M1, M2, R)) else 3
        if dt == 1:
            psi, L, M1, M2, R = [T.astype(1) for T in (psi, L, M1, M2, R)]
            mix = False
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
        X, c1, c2 = contract_meta(Lp, psi_m, dt, dev)
        ref_steps.append((plan_contraction(Lp, psi_m, X, c1, c2, dt), Lp, True, X))
        Y, c1, c2 = contract_meta(X, M1, dt, dev)
        ref_steps.append((plan_contraction(X, M1, Y, c1, c2, dt), M1, False, Y))
        Z, c1, c2 = contract_meta(Y, M2, dt, dev)
        ref_steps.append((plan_contraction(Y, M2, Z, c1, c2, dt), M2, False, Z))
        O, c1, c2 = contract_meta(Z, Rp, dt, dev)
        ref_steps.append((plan_contraction(Z, Rp, O, c1, c2, dt), Rp, False, O))
        if O.itoi() != psi.itoi() or [b.dim() for b in O.bonds()] != psi.shape():
            raise ValueError("Heff2: output block layout differs from psi's")
        self.ref_flops = float(sum(st[0].flops for st in ref_steps))
        self.pairs = [st[0].npairs for st in ref_steps]
        steps = ref_steps
        self.step_names = ["L.psi", ".M1", ".M2", ".R"]
        self.mode = "ref"
        if fuse_mpo:
            W12 = Contract(M1, M2)
            Zf, c1, c2 = contract_meta(X, W12, dt, dev)
            if Zf.labels() != Z.labels() or Zf.itoi() != Z.itoi():
                raise ValueError("Heff2: fused MPO step changed the intermediate layout")
            steps = [ref_steps[0], (plan_contraction(X, W12, Zf, c1, c2, dt), W12, False, Zf),
                     ref_steps[3]]
            self.step_names = ["L.psi", ".W12", ".R"]
            self.mode = "w12"
            if mix:
                try:
                    steps = self._mix_steps(Lp, psi_m, M1, M2, W12, Zf, Rp, psi, dev)
                    self.step_names = ["L.psi.W12 (mix)", ".R"]
                    self.mode = "mix"
                except MixUnsupported:
                    pass
        self.fused = bool(fuse_mpo)
        self.out_labels = O.labels()
        self.copies = [st[0].ws_bytes for st in steps]
        self.chain = _Chain(steps, dev)
        super().__init__(meta_elems(psi), dev)
        self._dtype = dt

    @staticmethod
    def _mix_steps(Lp, psi_m, M1, M2, W12, Zf, Rp, psi, dev):
        """The two launches of the mix chain: (mix plan, Lp) -> Y, then (Y . R) -> out."""
        lab_l, lab_w, lab_lp = Lp.labels()
        lab_r, lab_w2, lab_rp = Rp.labels()
        s1 = [x for x in psi.labels() if x in M1.labels()]
        s2 = [x for x in psi.labels() if x in M2.labels()]
        s1p = [x for x in M1.labels() if x not in (lab_w, s1[0]) and x not in M2.labels()]
        s2p = [x for x in M2.labels() if x not in (lab_w2, s2[0]) and x not in M1.labels()]
        if len(s1) != 1 or len(s2) != 1 or len(s1p) != 1 or len(s2p) != 1:
            raise MixUnsupported("Heff2 mix plan: unexpected network labels")
        # The labels of Y's two contracted legs sort as (w'', r), so the contraction over
        # (w'', r) reads Y and R (w'', r, r') as strided matrices.  Y's leg order is
        # (w'', r, l', s1', s2') when every block of it has one-dimensional physical legs: l'
        # is then contiguous, the contraction with R reads Y's rows (l') with unit stride, and
        # the mix launch computes transposed tiles (psi as its left operand).  Otherwise Y is
        # (l', s1', s2', w'', r) with r contiguous.  Either order gives the output legs
        # (l', s1', s2', r'), psi's layout.
        kw, kr = "~mix0", "~mix1"
        zb = {lab: Zf.bonds()[i] for i, lab in enumerate(Zf.labels())}
        Y = UniTensor([zb[lab_w2], zb[lab_r], zb[lab_lp], zb[s1p[0]], zb[s2p[0]]],
                      [kw, kr, lab_lp, s1p[0], s2p[0]], 2, 3, dev, _alloc=False)
        if Y.Nblocks() == 0 or np.any(Y._shapes[:, 3] * Y._shapes[:, 4] != 1):
            Y = UniTensor([zb[lab_lp], zb[s1p[0]], zb[s2p[0]], zb[lab_w2], zb[lab_r]],
                          [lab_lp, s1p[0], s2p[0], kw, kr], 3, 3, dev, _alloc=False)
        Wd = W12.to_dense()
        wl = W12.labels()
        Wd = np.ascontiguousarray(np.transpose(
            Wd, [wl.index(x) for x in (lab_w, s1[0], s1p[0], lab_w2, s2[0], s2p[0])]))
        args = (Lp, {"l": lab_l, "w": lab_w, "lp": lab_lp},
                psi_m, {"l": lab_l, "s1": s1[0], "s2": s2[0], "r": lab_r}, Wd)
        ylab = {"lp": lab_lp, "s1p": s1p[0], "s2p": s2p[0], "w2": kw, "r": kr}
        # Y blocks that no (w, s1, s2) slice reaches are zero for every psi, and Y blocks whose
        # (w'', r) sector pair no block of R has are never read: leave both out of Y, so
        # neither launch touches them
        Rpp = Rp.permute([lab_w2, lab_r, lab_rp]).relabel([kw, kr, lab_rp])
        rk = {(int(t[0]), int(t[1])) for t in Rpp._itoi}
        pw, pr = Y.labels().index(kw), Y.labels().index(kr)
        read = np.array([(int(t[pw]), int(t[pr])) in rk for t in Y._itoi], dtype=bool)
        keep = np.zeros(Y.Nblocks(), dtype=bool)
        arrays = mix_plan_arrays_native(*args, Y, ylab, keep, read)
        Y = subset_blocks(Y, keep)
        psi_left = psi_is_left(Y, ylab)
        na, nb = Lp._storage.numel(), meta_elems(psi_m)
        check_mix_bounds(arrays, *((nb, na) if psi_left else (na, nb)), meta_elems(Y))
        plan = _MixPlan(arrays, dev)
        O, c1, c2 = contract_meta(Y, Rpp, 3, dev)
        if O.itoi() != psi.itoi() or [b.dim() for b in O.bonds()] != psi.shape():
            raise ValueError("Heff2 mix chain: output block layout differs from psi's")
        return [(plan, Lp, not psi_left, Y),
                (plan_contraction(Y, Rpp, O, c1, c2, 3), Rpp, False, O)]

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


class _Flat:
    """A fixed chain operand given as a flat device tensor."""

    def __init__(self, t: torch.Tensor):
        self._storage = t.reshape(-1).contiguous()


def strided_batch_plan(nbatch: int, M: int, N: int, K: int, a: Sequence[int],
                       b: Sequence[int], c: Sequence[int], dtype: int = 3) -> _Plan:
    """cxa_strided_batch_plan: C_b = A_b . B_b for b < nbatch with A_b[m, k] at
    a[0]*b + a[1]*m + a[2]*k, B_b[k, n] at b[0]*b + b[1]*k + b[2]*n and C_b[m, n] at
    c[0]*b + c[1]*m + n (element offsets)."""
    h = ctypes.c_void_p(0)
    _lib.check(_lib.lib().cxa_strided_batch_plan(
        dtype, nbatch, M, N, K, a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1],
        ctypes.byref(h)))
    return _Plan(h.value, torch.cuda.current_device())


class Heff1Dense(_HeffBase):
    """Dense one-site L(l,w,l') A(l,s,r) W(w,w',s,s') R(r,w',r') with the order ((L.A).W).R,
    every step reading and writing its tensors in place (three launches, no permute copies):

        X (l', w, s, r)   = L'(l', w, l) . A(l, s, r)        over l, one GEMM
        Y (l', r, w', s') = X . W'(w, s, w', s')              over (w, s), a batch over r
        O (l', s', r')    = Y . R(r, w', r')                  over (r, w'), a batch over s'

    with L' and W' permuted once at construction (src/linalg/Tensordot.cpp:29-110 permutes
    both operands of every step into GEMM order instead).  Flops are the reference order's."""

    def __init__(self, L: torch.Tensor, W: torch.Tensor, R: torch.Tensor, A_shape):
        dev = L.device.index
        torch.cuda.set_device(dev)
        nl, D, nlp = (int(x) for x in L.shape)
        al, d, ar = (int(x) for x in A_shape)
        Dw, Dwp, ds, dsp = (int(x) for x in W.shape)
        rr, Dr, nrp = (int(x) for x in R.shape)
        if al != nl or Dw != D or ds != d or rr != ar or Dr != Dwp:
            raise ValueError("Heff1Dense: L %s, A %s, W %s and R %s do not match"
                             % (tuple(L.shape), tuple(A_shape), tuple(W.shape), tuple(R.shape)))
        self._Lp = _Flat(L.to(torch.float64).permute(2, 1, 0))      # (l', w, l)
        self._Wp = _Flat(W.to(torch.float64).permute(0, 2, 1, 3))   # (w, s, w', s')
        self._R = _Flat(R.to(torch.float64))                        # (r, w', r')
        p1 = strided_batch_plan(1, nlp * D, d * ar, nl, (0, nl, 1), (0, d * ar, 1),
                                (0, d * ar))
        p2 = strided_batch_plan(ar, nlp, Dwp * dsp, D * d, (1, D * d * ar, ar),
                                (0, Dwp * dsp, 1), (Dwp * dsp, ar * Dwp * dsp))
        p3 = strided_batch_plan(dsp, nlp, nrp, ar * Dwp, (1, ar * Dwp * dsp, dsp),
                                (0, nrp, 1), (nrp, dsp * nrp))
        steps = [(p1, self._Lp, True, nlp * D * d * ar),
                 (p2, self._Wp, False, nlp * ar * Dwp * dsp),
                 (p3, self._R, False, nlp * dsp * nrp)]
        self.psi_template = UniTensor([_trivial_bond(int(x), t) for x, t in
                                       zip(A_shape, [BD_KET, BD_BRA, BD_BRA])],
                                      ["l", "s", "r"], 1, 3, dev, _alloc=False)
        self.pairs = [st[0].npairs for st in steps]
        self.copies = [st[0].ws_bytes for st in steps]
        self.chain = _Chain(steps, dev)
        self.ref_flops = (2.0 * nlp * D * d * ar * nl + 2.0 * nlp * ar * Dwp * dsp * D * d
                          + 2.0 * nlp * dsp * nrp * ar * Dwp)
        super().__init__(nlp * dsp * nrp, dev)

    @property
    def flops(self) -> float:
        return self.ref_flops


__all__ = ["Heff2", "Heff1Dense", "dense_as_block"]
