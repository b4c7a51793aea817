"""U(1) Bond / UniTensor (dense and block) with the reference's labelled-contraction semantics.

Mirrors the parts of ``cytnx::Bond`` (include/Bond.hpp), ``cytnx::UniTensor``
(include/UniTensor.hpp), ``DenseUniTensor`` and ``BlockUniTensor`` (src/DenseUniTensor.cpp,
src/BlockUniTensor.cpp) that the contraction hot path uses.

A block tensor keeps all of its blocks in ONE flat device buffer, block b at element offset
``offsets[b]`` in row-major order, in the reference's block enumeration order.  Every block
contraction is one grouped GEMM launch planned by ``cxa_block_contract_plan``; plans are cached
by block-metadata signature so repeated contractions of the same shapes (Lanczos, sweeps) plan
once.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib, fileio
from .device import (CPU, TORCH_OF, Type_C128, Type_C64, Type_F32, Type_F64, dtype_id,
                     stream_handle, torch_device, type_promote)
from .tensor import Tensor

BD_KET, BD_BRA, BD_REG = -1, 1, 0
BD_IN, BD_OUT = BD_KET, BD_BRA

PREC_FP64, PREC_FP32, PREC_FP32X3 = 0, 1, 2


class Bond:
    """A bond: type (KET/BRA/REG) and, for symmetric bonds, U(1) qnums with degeneracies
    (include/Bond.hpp:31-148)."""

    def __init__(self, btype_or_dim, qnums=None, degs=None):
        if qnums is None:
            self._type = BD_REG
            self._qnums: List[int] = []
            self._degs: List[int] = []
            self._dim = int(btype_or_dim)
        else:
            self._type = int(btype_or_dim)
            if self._type == BD_REG:
                raise ValueError("a bond with qnums must be KET or BRA")
            self._qnums = [int(q[0]) if isinstance(q, (list, tuple)) else int(q) for q in qnums]
            self._degs = [int(d) for d in (degs if degs is not None else [1] * len(qnums))]
            if len(self._degs) != len(self._qnums):
                raise ValueError("qnums and degeneracies differ in length")
            self._dim = int(sum(self._degs))

    def type(self) -> int:
        return self._type

    def qnums(self) -> List[int]:
        return list(self._qnums)

    def getDegeneracies(self) -> List[int]:
        return list(self._degs)

    def dim(self) -> int:
        return self._dim

    def has_symmetry(self) -> bool:
        return bool(self._qnums)

    def redirect(self) -> "Bond":
        b = self.clone()
        b._type = -b._type
        return b

    def clone(self) -> "Bond":
        b = Bond.__new__(Bond)
        b._type, b._qnums, b._degs, b._dim = self._type, list(self._qnums), list(self._degs), \
            self._dim
        return b

    def __eq__(self, other) -> bool:
        return (isinstance(other, Bond) and self._type == other._type and
                self._qnums == other._qnums and self._degs == other._degs and
                self._dim == other._dim)

    def __repr__(self) -> str:
        if not self._qnums:
            return "Bond(dim=%d)" % self._dim
        return "Bond(%s, qnums=%s, degs=%s)" % ("KET" if self._type == BD_KET else "BRA",
                                               self._qnums, self._degs)


def enumerate_blocks(bonds: Sequence[Bond]) -> np.ndarray:
    """Zero-flux qindex tuples (nblocks x rank) in the reference's order, via the C++ walker."""
    rank = len(bonds)
    if rank == 0:
        return np.zeros((1, 0), dtype=np.uint64)
    L = _lib.lib()
    btype = _lib.int_array([b.type() for b in bonds])
    nq = _lib.i64_array([len(b.qnums()) for b in bonds])
    qn = _lib.i64_array([q for b in bonds for q in b.qnums()])
    cap = 4096
    while True:
        out = np.zeros((cap, rank), dtype=np.uint64)
        n = ctypes.c_int64(0)
        _lib.check(L.cxa_u1_enumerate_blocks(rank, btype, nq, qn, cap,
                                             out.ctypes.data_as(_lib.u64p), ctypes.byref(n)))
        if n.value <= cap:
            return out[:n.value].copy()
        cap = n.value


class _Plan:
    """Owns a cxa_bc_plan handle and its workspace."""

    def __init__(self, handle: int, device: int):
        self.handle = handle
        self.device = device
        f = ctypes.c_double(0)
        ws = ctypes.c_int64(0)
        npairs = ctypes.c_int64(0)
        _lib.check(_lib.lib().cxa_block_contract_info(handle, ctypes.byref(f), ctypes.byref(ws),
                                                      ctypes.byref(npairs)))
        self.flops = f.value
        self.ws_bytes = ws.value
        self.npairs = npairs.value
        self._ws = None

    def workspace(self) -> int:
        if self.ws_bytes == 0:
            return 0
        if self._ws is None:
            self._ws = torch.empty(self.ws_bytes, dtype=torch.uint8,
                                   device=torch_device(self.device))
        return self._ws.data_ptr()

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().cxa_block_contract_destroy(self.handle)
        except Exception:
            pass


_PLAN_CACHE: Dict[tuple, _Plan] = {}


def _plan_for(dtype: int, A: "UniTensor", B: "UniTensor", c1: List[int], c2: List[int],
              C: "UniTensor") -> _Plan:
    key = (dtype, A.device(), A._itoi.tobytes(), A._shapes.tobytes(), A._offsets.tobytes(),
           B._itoi.tobytes(), B._shapes.tobytes(), B._offsets.tobytes(), tuple(c1), tuple(c2),
           C._itoi.tobytes(), C._shapes.tobytes(), C._offsets.tobytes(), A.rank(), B.rank())
    p = _PLAN_CACHE.get(key)
    if p is not None:
        return p
    L = _lib.lib()
    h = ctypes.c_void_p(0)
    torch.cuda.set_device(A.device())
    _lib.check(L.cxa_block_contract_plan(
        dtype,
        A.rank(), A.Nblocks(), A._itoi.ctypes.data_as(_lib.u64p),
        A._shapes.ctypes.data_as(_lib.u64p), A._offsets.ctypes.data_as(_lib.i64p),
        B.rank(), B.Nblocks(), B._itoi.ctypes.data_as(_lib.u64p),
        B._shapes.ctypes.data_as(_lib.u64p), B._offsets.ctypes.data_as(_lib.i64p),
        len(c1), _lib.u64_array(c1), _lib.u64_array(c2),
        C.Nblocks(), C._itoi.ctypes.data_as(_lib.u64p), C._shapes.ctypes.data_as(_lib.u64p),
        C._offsets.ctypes.data_as(_lib.i64p), ctypes.byref(h)))
    p = _Plan(h.value, A.device())
    if len(_PLAN_CACHE) > 4096:
        _PLAN_CACHE.clear()
    _PLAN_CACHE[key] = p
    return p


class _BPPlan:
    """A block permute plan (cxa_block_permute_plan) and its destruction."""

    def __init__(self, handle: int):
        self.handle = handle

    def __del__(self):
        try:
            if self.handle:
                _lib.lib().cxa_block_permute_destroy(self.handle)
        except Exception:
            pass


_BP_CACHE: Dict[tuple, _BPPlan] = {}


def block_permute_plan(dtype: int, device: int, shapes: np.ndarray, offsets: np.ndarray,
                       mapper: Sequence[int]) -> _BPPlan:
    """The (cached) plan that permutes every block of a layout (row-major blocks of extents
    shapes[b] at element offsets offsets[b]) by mapper into the same offsets."""
    shp = np.ascontiguousarray(shapes, dtype=np.uint64)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    rank = shp.shape[1] if shp.ndim == 2 else 0
    key = (dtype, device, rank, shp.tobytes(), off.tobytes(), tuple(int(m) for m in mapper))
    p = _BP_CACHE.get(key)
    if p is not None:
        return p
    h = ctypes.c_void_p(0)
    torch.cuda.set_device(device)
    _lib.check(_lib.lib().cxa_block_permute_plan(
        dtype, rank, off.size, shp.ctypes.data_as(_lib.u64p), off.ctypes.data_as(_lib.i64p),
        off.ctypes.data_as(_lib.i64p), _lib.u64_array(list(mapper)), ctypes.byref(h)))
    p = _BPPlan(h.value)
    if len(_BP_CACHE) > 4096:
        _BP_CACHE.clear()
    _BP_CACHE[key] = p
    return p


def _block_layout(bonds: Sequence[Bond], itoi: np.ndarray) -> Tuple[np.ndarray, np.ndarray, int]:
    rank = len(bonds)
    degs = [np.asarray(b.getDegeneracies(), dtype=np.uint64) for b in bonds]
    nb = itoi.shape[0]
    shapes = np.zeros((nb, rank), dtype=np.uint64)
    for i in range(rank):
        shapes[:, i] = degs[i][itoi[:, i].astype(np.int64)]
    sizes = np.prod(shapes.astype(np.int64), axis=1) if rank else np.ones(nb, dtype=np.int64)
    offsets = np.zeros(nb, dtype=np.int64)
    if nb:
        offsets[1:] = np.cumsum(sizes)[:-1]
    total = int(sizes.sum()) if nb else 0
    return shapes, offsets, total


class UniTensor:
    """Dense or U(1)-block UniTensor.

    Block tensors: ``_bonds`` (all KET/BRA), ``_itoi`` (nblocks x rank qindices), ``_shapes``,
    ``_offsets`` and one flat ``_storage`` buffer.  Dense tensors wrap one ``Tensor`` block.
    """

    def __init__(self, bonds: Sequence[Bond] = (), labels: Optional[Sequence[str]] = None,
                 rowrank: int = -1, dtype: int = Type_F64, device: int = CPU,
                 is_diag: bool = False, name: str = "", _alloc: bool = True):
        self._is_diag = bool(is_diag)
        if self._is_diag:
            self._init_diag(bonds, labels, rowrank, dtype, device, name, _alloc)
            return
        self._bonds = [b.clone() for b in bonds]
        self._labels = [str(x) for x in labels] if labels is not None else \
            [str(i) for i in range(len(bonds))]
        if len(self._labels) != len(self._bonds):
            raise ValueError("labels must have same length as # of bonds.")
        if len(set(self._labels)) != len(self._labels):
            raise ValueError("labels cannot contain duplicated elements.")
        self._name = name
        self._is_block = bool(self._bonds) and all(b.has_symmetry() for b in self._bonds)
        if any(b.has_symmetry() for b in self._bonds) and not self._is_block:
            raise ValueError("All bonds must be tagged for UniTensor with symmetries.")
        if self._is_block:
            nket = sum(1 for b in self._bonds if b.type() == BD_KET)
            self._rowrank = nket if rowrank == -1 else int(rowrank)
            self._itoi = enumerate_blocks(self._bonds)
            self._shapes, self._offsets, total = _block_layout(self._bonds, self._itoi)
            self._storage = torch.zeros(total if _alloc else 0, dtype=TORCH_OF[dtype],
                                        device=torch_device(device))
            self._block = None
        else:
            self._rowrank = (len(self._bonds) + 1) // 2 if rowrank == -1 else int(rowrank)
            self._block = Tensor([b.dim() for b in self._bonds], dtype, device)
        if self._rowrank < 0 or self._rowrank > len(self._bonds):
            raise ValueError("rowrank must be >=0 and <=rank.")

    def _init_diag(self, bonds, labels, rowrank, dtype, device, name, alloc) -> None:
        """Diagonal rank-2 UniTensor (the singular values of an SVD): only the diagonal is
        stored.  Block form: bonds (b, b.redirect()), block i is the diagonal of the (i, i)
        sector, deg_i elements, in qnum order (UniTensor.hpp is_diag, src/BlockUniTensor.cpp
        Init with is_diag)."""
        if len(bonds) != 2:
            raise ValueError("[ERROR] is_diag=True requires a rank-2 UniTensor.")
        b0, b1 = bonds[0], bonds[1]
        if b0.dim() != b1.dim():
            raise ValueError("[ERROR] is_diag=True requires the two bonds to have the same "
                             "dimension.")
        self._bonds = [b0.clone(), b1.clone()]
        self._labels = [str(x) for x in labels] if labels is not None else ["0", "1"]
        if len(self._labels) != 2 or self._labels[0] == self._labels[1]:
            raise ValueError("invalid labels for a diagonal UniTensor")
        self._name = name
        self._rowrank = 1 if rowrank == -1 else int(rowrank)
        self._is_block = b0.has_symmetry()
        if self._is_block:
            if b0.qnums() != b1.qnums() or b0.getDegeneracies() != b1.getDegeneracies() or \
                    b0.type() + b1.type() != 0:
                raise ValueError("a diagonal block UniTensor needs bonds (b, b.redirect())")
            nq = len(b0.qnums())
            self._itoi = np.repeat(np.arange(nq, dtype=np.uint64)[:, None], 2, axis=1)
            degs = np.asarray(b0.getDegeneracies(), dtype=np.uint64)
            self._shapes = np.stack([degs, degs], axis=1)
            self._offsets = np.zeros(nq, dtype=np.int64)
            if nq:
                self._offsets[1:] = np.cumsum(degs.astype(np.int64))[:-1]
            total = int(degs.sum())
            self._storage = torch.zeros(total if alloc else 0, dtype=TORCH_OF[dtype],
                                        device=torch_device(device))
            self._block = None
        else:
            self._block = Tensor([b0.dim()], dtype, device)

    def is_diag(self) -> bool:
        return getattr(self, "_is_diag", False)

    # ------------------------------------------------------------------ metadata
    def is_blockform(self) -> bool:
        return self._is_block

    def rank(self) -> int:
        return len(self._bonds)

    def rowrank(self) -> int:
        return self._rowrank

    def labels(self) -> List[str]:
        return list(self._labels)

    def bonds(self) -> List[Bond]:
        return [b.clone() for b in self._bonds]

    def shape(self) -> List[int]:
        return [b.dim() for b in self._bonds]

    def dtype(self) -> int:
        return dtype_id(self._storage.dtype) if self._is_block else self._block.dtype()

    def device(self) -> int:
        s = self._storage if self._is_block else self._block.storage()
        return CPU if s.device.type == "cpu" else s.device.index

    def Nblocks(self) -> int:
        return int(self._itoi.shape[0]) if self._is_block else 1

    def set_labels(self, labels: Sequence[str]) -> "UniTensor":
        labels = [str(x) for x in labels]
        if len(labels) != self.rank() or len(set(labels)) != len(labels):
            raise ValueError("invalid labels")
        self._labels = labels
        return self

    relabel_ = set_labels

    def relabel(self, labels: Sequence[str]) -> "UniTensor":
        return self.clone_meta().set_labels(labels)

    def set_rowrank_(self, r: int) -> "UniTensor":
        if r < 0 or r > self.rank():
            raise ValueError("rowrank must be >=0 and <=rank.")
        self._rowrank = int(r)
        return self

    def set_name_(self, name: str) -> "UniTensor":
        self._name = name
        return self

    def name(self) -> str:
        return self._name

    def storage(self) -> torch.Tensor:
        return self._storage if self._is_block else self._block.storage()

    # ------------------------------------------------------------------ blocks
    def get_block_(self, i: int = 0) -> torch.Tensor:
        """View of block i (torch, in place on the flat buffer)."""
        if not self._is_block:
            return self._block.torch()
        o = int(self._offsets[i])
        if self.is_diag():
            return self._storage[o:o + int(self._shapes[i][0])]
        n = int(np.prod(self._shapes[i].astype(np.int64))) if self.rank() else 1
        return self._storage[o:o + n].view([int(s) for s in self._shapes[i]])

    def get_blocks_(self) -> List[torch.Tensor]:
        return [self.get_block_(i) for i in range(self.Nblocks())]

    def put_block(self, data, i: int = 0) -> "UniTensor":
        blk = self.get_block_(i)
        src = torch.as_tensor(np.asarray(data) if not isinstance(data, torch.Tensor) else data)
        if list(src.shape) != list(blk.shape):
            raise ValueError("block shape mismatch")
        blk.copy_(src.to(blk.device, blk.dtype))
        return self

    def itoi(self) -> List[List[int]]:
        return self._itoi.astype(np.int64).tolist()

    def clone_meta(self) -> "UniTensor":
        out = UniTensor.__new__(UniTensor)
        out.__dict__.update(self.__dict__)
        out._bonds = [b.clone() for b in self._bonds]
        out._labels = list(self._labels)
        return out

    def clone(self) -> "UniTensor":
        out = self.clone_meta()
        if self._is_block:
            out._storage = self._storage.clone()
        else:
            out._block = self._block.clone()
        return out

    def to(self, device: int) -> "UniTensor":
        out = self.clone_meta()
        if self._is_block:
            out._storage = self._storage.to(torch_device(device))
        else:
            out._block = self._block.to(device)
        return out

    def astype(self, dtype: int) -> "UniTensor":
        if dtype == self.dtype():
            return self
        out = self.clone_meta()
        if self._is_block:
            out._storage = self._storage.to(TORCH_OF[dtype])
        else:
            t = self._block.contiguous()
            out._block = Tensor.from_storage(t.storage().to(TORCH_OF[dtype]), t.shape())
        return out

    def contiguous_(self) -> "UniTensor":
        if not self._is_block:
            self._block.contiguous_()
        return self

    # ------------------------------------------------------------------ host views
    def to_dense(self) -> np.ndarray:
        """Dense host array (zeros outside the blocks)."""
        if not self._is_block:
            if self.is_diag():
                return np.diag(self._block.numpy())
            return self._block.numpy()
        out = np.zeros(self.shape(), dtype=fileio.NP_DTYPE[self.dtype()])
        offs = [np.concatenate([[0], np.cumsum(b.getDegeneracies())[:-1]]).astype(int)
                for b in self._bonds]
        host = self._storage.detach().cpu().numpy()
        if self.is_diag():
            for bi in range(self.Nblocks()):
                d = int(self._shapes[bi][0])
                o = int(self._offsets[bi])
                s0 = offs[0][int(self._itoi[bi][0])]
                out[np.arange(s0, s0 + d), np.arange(s0, s0 + d)] = host[o:o + d]
            return out
        for bi in range(self.Nblocks()):
            shp = [int(s) for s in self._shapes[bi]]
            o = int(self._offsets[bi])
            blk = host[o:o + int(np.prod(shp))].reshape(shp)
            sl = tuple(slice(offs[a][int(q)], offs[a][int(q)] + shp[a])
                       for a, q in enumerate(self._itoi[bi]))
            out[sl] = blk
        return out

    def item(self):
        if self.rank() != 0:
            raise ValueError("item() requires a rank-0 UniTensor")
        return self.storage().reshape(-1)[0].item()

    # ------------------------------------------------------------------ construction helpers
    @classmethod
    def from_saved(cls, ut: fileio.SavedUniTensor, device: int = CPU) -> "UniTensor":
        if ut.uten_type == fileio.UTEN_DENSE:
            bonds = [Bond(b.dim) for b in ut.bonds]
            out = cls(bonds, ut.labels, ut.rowrank, ut.blocks[0].dtype, device,
                      is_diag=ut.is_diag, name=ut.name)
            out._block = Tensor.from_numpy(ut.blocks[0].logical(), device)
            return out
        if ut.uten_type != fileio.UTEN_BLOCK:
            raise NotImplementedError("only dense and block UniTensors are supported")
        bonds = [Bond(b.btype, [q[0] for q in b.qnums], b.degs) for b in ut.bonds]
        dtype = ut.blocks[0].dtype if ut.blocks else Type_F64
        out = cls(bonds, ut.labels, ut.rowrank, dtype, device, is_diag=ut.is_diag, name=ut.name)
        want = {tuple(int(x) for x in t): i for i, t in enumerate(out._itoi)}
        for t, blk in zip(ut.itoi, ut.blocks):
            out.put_block(blk.logical(), want[tuple(t)])
        return out

    @classmethod
    def Load(cls, path: str, device: int = CPU) -> "UniTensor":
        return cls.from_saved(fileio.read_unitensor(path), device)

    def _is_braket_form(self) -> bool:
        """UniTensor_base::_update_braket (include/UniTensor.hpp:70-86)."""
        if not self._bonds or self._bonds[0].type() == BD_REG:
            return False
        return all(b.type() == (BD_KET if i < self._rowrank else BD_BRA)
                   for i, b in enumerate(self._bonds))

    def to_saved(self) -> fileio.SavedUniTensor:
        """The file content of this UniTensor (blocks in their stored order)."""
        u1 = [fileio.SavedSymmetry(fileio.SYM_U1, 1)]
        if self._is_block:
            bonds = [fileio.SavedBond(b.type(), [[q] for q in b.qnums()], b.getDegeneracies(),
                                      list(u1), b.dim()) for b in self._bonds]
            ut = fileio.SavedUniTensor(fileio.UTEN_BLOCK, self._is_braket_form(), True,
                                       self.is_diag(), self._rowrank, self._name,
                                       list(self._labels), bonds)
            ut.itoi = self._itoi.astype(np.int64).tolist() if self.rank() else \
                [[] for _ in range(self.Nblocks())]
            ut.blocks = [self.get_block_(i).detach().cpu().numpy()
                         for i in range(self.Nblocks())]
            if not self._bonds:
                ut.syms_cache = list(u1)
            return ut
        bonds = [fileio.SavedBond(b.type(), [], [], [], b.dim()) for b in self._bonds]
        tag = bool(self._bonds) and self._bonds[0].type() != BD_REG
        ut = fileio.SavedUniTensor(fileio.UTEN_DENSE, self._is_braket_form(), tag,
                                   self.is_diag(), self._rowrank, self._name,
                                   list(self._labels), bonds)
        ut.blocks = [self._block.numpy()]
        return ut

    def Save(self, fname: str) -> None:
        """UniTensor::Save (src/UniTensor.cpp:169-187): '.cytnx' is appended to a name without
        an extension."""
        import os
        if not os.path.splitext(fname)[1]:
            fname = fname + ".cytnx"
        fileio.write_unitensor(fname, self.to_saved())

    # ------------------------------------------------------------------ permute
    def _resolve_mapper(self, mapper) -> List[int]:
        mapper = list(mapper)
        if mapper and isinstance(mapper[0], str):
            return [self._labels.index(m) for m in mapper]
        return [int(m) for m in mapper]

    def permute(self, mapper, rowrank: int = -1) -> "UniTensor":
        """UniTensor::permute: permutes bonds, labels and (block) itoi; the data of every block
        is permuted on the GPU (BlockUniTensor::permute, src/BlockUniTensor.cpp:565-633, then
        contiguous)."""
        mp = self._resolve_mapper(mapper)
        if sorted(mp) != list(range(self.rank())):
            raise ValueError("invalid permutation")
        out = self.clone_meta()
        out._bonds = [self._bonds[m].clone() for m in mp]
        out._labels = [self._labels[m] for m in mp]
        out._rowrank = self._rowrank if rowrank < 0 else int(rowrank)
        if out._rowrank < 0 or out._rowrank > self.rank():
            raise ValueError("rowrank must be >=0 and <=rank.")
        if self.is_diag():
            # a diagonal stays diagonal: only the bonds and labels trade places
            return out
        if not self._is_block:
            out._block = self._block.permute(mp)
            return out
        out._itoi = np.ascontiguousarray(self._itoi[:, mp])
        out._shapes = np.ascontiguousarray(self._shapes[:, mp])
        dst = torch.empty_like(self._storage)
        L = _lib.lib()
        if self._storage.numel():
            if self._storage.device.type == "cpu":
                raise NotImplementedError("cytnx_amd implements the Device.cuda branch")
            plan = block_permute_plan(self.dtype(), self.device(), self._shapes, self._offsets,
                                      mp)
            _lib.check(L.cxa_block_permute_exec(plan.handle, self._storage.data_ptr(),
                                                dst.data_ptr(), stream_handle(self.device())))
        out._storage = dst
        return out

    def permute_(self, mapper, rowrank: int = -1) -> "UniTensor":
        p = self.permute(mapper, rowrank)
        self.__dict__.update(p.__dict__)
        return self

    # ------------------------------------------------------------------ conjugation
    def Conj(self) -> "UniTensor":
        """Complex conjugate (real dtypes: the same data)."""
        out = self.clone_meta()
        if self._is_block:
            st = self._storage
            out._storage = torch.conj_physical(st) if st.is_complex() else st
        else:
            t = self._block.contiguous()
            st = t.storage()
            out._block = Tensor.from_storage(torch.conj_physical(st) if st.is_complex() else st,
                                             t.shape())
        return out

    def Conj_(self) -> "UniTensor":
        self.__dict__.update(self.Conj().__dict__)
        return self

    def Transpose(self) -> "UniTensor":
        """Every bond redirected, the index order reversed, rowrank -> rank - rowrank
        (BlockUniTensor::Transpose_, src/BlockUniTensor.cpp:1237-1245)."""
        rank = self.rank()
        out = self.clone_meta()
        out._bonds = [b.redirect() if b.type() != BD_REG else b.clone() for b in self._bonds]
        return out.permute(list(range(rank - 1, -1, -1)), rank - self._rowrank)

    def Transpose_(self) -> "UniTensor":
        self.__dict__.update(self.Transpose().__dict__)
        return self

    def Dagger(self) -> "UniTensor":
        """Conj().Transpose() (include/UniTensor.hpp:1731-1738)."""
        return self.Conj().Transpose()

    def Dagger_(self) -> "UniTensor":
        self.__dict__.update(self.Dagger().__dict__)
        return self

    # ------------------------------------------------------------------ arithmetic
    def _with_storage(self, st: torch.Tensor) -> "UniTensor":
        out = self.clone_meta()
        if self._is_block:
            out._storage = st
        else:
            out._block = Tensor.from_storage(st, self._block.contiguous().shape())
        return out

    def _flat(self) -> torch.Tensor:
        return self._storage if self._is_block else self._block.contiguous().storage()

    def _same_layout(self, other: "UniTensor") -> bool:
        if self._is_block != other._is_block or self.is_diag() != other.is_diag():
            return False
        if self._is_block:
            return (self._itoi.shape == other._itoi.shape and
                    bool(np.all(self._itoi == other._itoi)) and
                    bool(np.all(self._shapes == other._shapes)))
        return self.shape() == other.shape()

    def __mul__(self, other):
        if isinstance(other, UniTensor):
            if not self._same_layout(other):
                raise ValueError("elementwise product needs UniTensors of the same layout")
            return self._with_storage(self._flat() * other._flat())
        return self._with_storage(self._flat() * other)

    __rmul__ = __mul__

    def __truediv__(self, other):
        if isinstance(other, UniTensor):
            if not self._same_layout(other):
                raise ValueError("elementwise division needs UniTensors of the same layout")
            return self._with_storage(self._flat() / other._flat())
        return self._with_storage(self._flat() / other)

    def __neg__(self):
        return self._with_storage(-self._flat())

    def __add__(self, other: "UniTensor"):
        if not isinstance(other, UniTensor) or not self._same_layout(other):
            raise ValueError("addition needs UniTensors of the same layout")
        return self._with_storage(self._flat() + other._flat())

    def __sub__(self, other: "UniTensor"):
        if not isinstance(other, UniTensor) or not self._same_layout(other):
            raise ValueError("subtraction needs UniTensors of the same layout")
        return self._with_storage(self._flat() - other._flat())

    def normalize(self) -> "UniTensor":
        return self / self.Norm()

    def normalize_(self) -> "UniTensor":
        self.__dict__.update(self.normalize().__dict__)
        return self

    # ------------------------------------------------------------------ norms
    def Norm(self) -> float:
        """2-norm of all elements (BlockUniTensor::Norm, src/BlockUniTensor.cpp:1356-1372: the
        blocks together), by cxa_nrm2 on the device."""
        from .linalg import norm
        if self.storage().numel() == 0:
            return 0.0
        return norm(self)

    norm = Norm

    # ------------------------------------------------------------------ contraction
    def contract(self, rhs: "UniTensor", mv_elem_self: bool = False,
                 mv_elem_rhs: bool = False) -> "UniTensor":
        return Contract(self, rhs)

    def __repr__(self) -> str:
        kind = "Block" if self._is_block else "Dense"
        return "UniTensor<%s>(labels=%s, shape=%s, rowrank=%d, nblocks=%d)" % (
            kind, self._labels, self.shape(), self._rowrank, self.Nblocks())


def _common(la: Sequence[str], lb: Sequence[str]):
    """vec_intersect_ (src/utils/vec_intersect.cpp:35-51): sorted common labels."""
    comm = sorted(set(la) & set(lb))
    return comm, [list(la).index(c) for c in comm], [list(lb).index(c) for c in comm]


def _promote_pair(A: UniTensor, B: UniTensor) -> Tuple[UniTensor, UniTensor, int]:
    dt = type_promote(A.dtype(), B.dtype())
    return A.astype(dt), B.astype(dt), dt


def contract_meta(A: UniTensor, B: UniTensor, dt: Optional[int] = None,
                  device: Optional[int] = None):
    """Output metadata of A.contract(B) without data: (out, c1, c2).

    Common labels sorted (vec_intersect_), output bonds/labels = L free then R free
    (src/BlockUniTensor.cpp:948-968), rowrank arithmetic of :964-968, output blocks enumerated
    by the zero-flux rule.  ``out`` is a block UniTensor with an empty storage."""
    comm, c1, c2 = _common(A.labels(), B.labels())
    for i, j in zip(c1, c2):
        if A.is_blockform() and A._bonds[i].type() + B._bonds[j].type() != 0:
            raise ValueError("A bond with type BRA can only be contracted with type KET. "
                             "Invalid contraction for label '%s'." % A._labels[i])
        if A._bonds[i].dim() != B._bonds[j].dim():
            raise ValueError("contracted bond dimensions differ for label '%s'" % A._labels[i])
    nc1 = [i for i in range(A.rank()) if i not in c1]
    nc2 = [i for i in range(B.rank()) if i not in c2]
    out_bonds = [A._bonds[i] for i in nc1] + [B._bonds[i] for i in nc2]
    out_labels = [A._labels[i] for i in nc1] + [B._labels[i] for i in nc2]
    rr = A.rowrank() + B.rowrank() - sum(1 for i in c1 if i < A.rowrank()) \
        - sum(1 for i in c2 if i < B.rowrank())
    dt = A.dtype() if dt is None else dt
    dev = A.device() if device is None else device
    if out_bonds:
        out = UniTensor(out_bonds, out_labels, max(0, min(rr, len(out_bonds))), dt, dev,
                        _alloc=False)
    else:
        out = _scalar_block(dt, dev)
        out._storage = out._storage[:0]
    return out, c1, c2


def meta_elems(T: UniTensor) -> int:
    """Number of elements of all blocks together."""
    if T.rank() == 0:
        return 1
    return int(np.sum(np.prod(T._shapes.astype(np.int64), axis=1))) if T.Nblocks() else 0


def plan_contraction(A: UniTensor, B: UniTensor, C: UniTensor, c1, c2, dtype: int) -> "_Plan":
    """Cached cxa_block_contract_plan for A x B -> C (metadata only)."""
    return _plan_for(dtype, A, B, list(c1), list(c2), C)


def Contract(A, B=None, *args, **kwargs) -> UniTensor:
    """The two overloads of the Python binding (pybind/unitensor_py.cpp:2655-2666):

    * ``Contract(Tl, Tr, cacheL=False, cacheR=False)``: UniTensor::contract on the GPU
      (src/BlockUniTensor.cpp:777-1235, src/DenseUniTensor.cpp:1022-1164); the cache flags are
      accepted and have no effect (plans are cached by metadata anyway);
    * ``Contract(TNs, order="", optimal=True)``: a list of UniTensors, as ``Contracts``
      (src/UniTensor_base.cpp:711-714)."""
    if isinstance(A, (list, tuple)):
        from .network import Contracts
        return Contracts(A, *(([] if B is None else [B]) + list(args)), **kwargs)
    for k in kwargs:
        if k not in ("cacheL", "cacheR"):
            raise TypeError("Contract() got an unexpected keyword argument '%s'" % k)
    if len(args) > 2 or B is None:
        raise TypeError("Contract(Tl, Tr, cacheL=False, cacheR=False)")
    return _contract_pair(A, B)


def _contract_pair(A: UniTensor, B: UniTensor) -> UniTensor:
    """A.contract(B): common labels contracted, free legs of A then B."""
    if A.is_blockform() != B.is_blockform():
        raise TypeError("Cannot contract BlockUniTensor with other type of UniTensor.")
    comm, c1, c2 = _common(A.labels(), B.labels())
    for i, j in zip(c1, c2):
        if A.is_blockform() and A._bonds[i].type() + B._bonds[j].type() != 0:
            raise ValueError("A bond with type BRA can only be contracted with type KET. "
                             "Invalid contraction for label '%s'." % A._labels[i])
        if A._bonds[i].dim() != B._bonds[j].dim():
            raise ValueError("contracted bond dimensions differ for label '%s'" % A._labels[i])
    A, B, dt = _promote_pair(A, B)
    nc1 = [i for i in range(A.rank()) if i not in c1]
    nc2 = [i for i in range(B.rank()) if i not in c2]
    out_bonds = [A._bonds[i] for i in nc1] + [B._bonds[i] for i in nc2]
    out_labels = [A._labels[i] for i in nc1] + [B._labels[i] for i in nc2]
    rr = A.rowrank() + B.rowrank() - sum(1 for i in c1 if i < A.rowrank()) \
        - sum(1 for i in c2 if i < B.rowrank())
    if A.is_diag() or B.is_diag():
        return _contract_diag(A, B, c1, c2, out_labels, max(0, min(rr, len(out_bonds))))
    if not A.is_blockform():
        from . import linalg
        if not comm:
            raise NotImplementedError("dense outer product (no common label) is not on the path")
        T = linalg.Tensordot(A._block, B._block, c1, c2)
        out = UniTensor.__new__(UniTensor)
        out._bonds = [b.clone() for b in out_bonds]
        out._labels = out_labels
        out._name = ""
        out._is_block = False
        out._rowrank = rr
        out._block = T
        return out
    dev = A.device()
    if dev == CPU or B.device() != dev:
        raise NotImplementedError("block contraction runs on one GPU; move both operands there")
    out = UniTensor(out_bonds, out_labels, max(0, min(rr, len(out_bonds))), dt, dev,
                    _alloc=False) if out_bonds else _scalar_block(dt, dev)
    out._storage = torch.empty(int(np.sum(np.prod(out._shapes.astype(np.int64), axis=1)))
                               if out.Nblocks() and out.rank() else max(1, out.Nblocks()),
                               dtype=TORCH_OF[dt], device=torch_device(dev))
    plan = _plan_for(dt, A, B, c1, c2, out)
    _lib.check(_lib.lib().cxa_block_contract_exec(
        plan.handle, A._storage.data_ptr(), B._storage.data_ptr(), out._storage.data_ptr(),
        plan.workspace(), PREC_FP64, stream_handle(dev)))
    return out


def _contract_diag(A: UniTensor, B: UniTensor, c1: List[int], c2: List[int],
                   out_labels: List[str], rowrank: int) -> UniTensor:
    """Contraction with a diagonal operand over one common leg: the other operand's blocks
    scaled along that leg (Tensordot_dg, src/linalg/Tensordot_dg.cpp; the block case of
    src/BlockUniTensor.cpp:1026-1044), then the legs ordered as Contract orders them."""
    if len(c1) != 1:
        raise NotImplementedError("contraction with a diagonal UniTensor over %d legs"
                                  % len(c1))
    if A.is_diag() and B.is_diag():
        out = A.clone_meta()
        out._bonds = [A._bonds[1 - c1[0]].clone(), B._bonds[1 - c2[0]].clone()]
        out._labels = list(out_labels)
        out._rowrank = 1
        if A._is_block:
            out._storage = A._storage * B._storage
        else:
            out._block = Tensor.from_storage(A._block.storage() * B._block.storage(),
                                             [A._bonds[0].dim()])
        return out
    diag_left = A.is_diag()
    D, T, ax, dax = (A, B, c2[0], c1[0]) if diag_left else (B, A, c1[0], c2[0])
    if T.device() == CPU:
        raise NotImplementedError("block contraction runs on one GPU; move both operands there")
    if T._is_block and (T._bonds[ax].qnums() != D._bonds[dax].qnums() or
                        T._bonds[ax].getDegeneracies() != D._bonds[dax].getDegeneracies()):
        raise ValueError("the contracted bond of the diagonal UniTensor differs from its partner")
    out = T.clone_meta()
    out._bonds = list(T._bonds)
    out._bonds[ax] = D._bonds[1 - dax].clone()
    out._labels = list(T._labels)
    out._labels[ax] = D._labels[1 - dax]
    if T._is_block:
        st = torch.empty_like(T._storage)
        for bi in range(T.Nblocks()):
            q = int(T._itoi[bi][ax])
            d = D.get_block_(q)
            shp = [1] * T.rank()
            shp[ax] = -1
            blk = T.get_block_(bi)
            o = int(T._offsets[bi])
            st[o:o + blk.numel()].view(blk.shape).copy_(blk * d.view(shp))
        out._storage = st
    else:
        t = T._block.contiguous()
        shp = [1] * T.rank()
        shp[ax] = -1
        v = t.torch() * D._block.torch().view(shp)
        out._block = Tensor.from_storage(v.reshape(-1).contiguous(), t.shape())
    order = [out._labels.index(l) for l in out_labels]
    if order != list(range(out.rank())):
        out = out.permute(order, rowrank)
    out._rowrank = rowrank
    return out


def _scalar_block(dt: int, dev: int) -> UniTensor:
    out = UniTensor.__new__(UniTensor)
    out._bonds = []
    out._labels = []
    out._name = ""
    out._is_block = True
    out._rowrank = 0
    out._itoi = np.zeros((1, 0), dtype=np.uint64)
    out._shapes = np.zeros((1, 0), dtype=np.uint64)
    out._offsets = np.zeros(1, dtype=np.int64)
    out._block = None
    out._storage = torch.zeros(1, dtype=TORCH_OF[dt], device=torch_device(dev))
    return out


__all__ = ["Bond", "UniTensor", "Contract", "BD_KET", "BD_BRA", "BD_REG", "BD_IN", "BD_OUT",
           "enumerate_blocks"]
