"""Reader for Cytnx's binary ``.cytn`` (Tensor) and ``.cytnx`` (UniTensor) files.

The layout is little-endian with native ``sizeof`` and is defined by the reference's
``_Save``/``_Load`` pairs:

* Storage   -- ``src/backend/Storage.cpp:161-193`` (``u32 999, u64 size, u32 dtype, i32 device``,
  raw element bytes).
* Tensor    -- ``src/Tensor.cpp:472-551`` (``u32 889, u32 version`` or legacy ``u32 888``,
  ``u64 rank, u64 contiguous, u64 shape[], u64 mapper[], u64 invmapper[]``, Storage).
* Symmetry  -- ``src/Symmetry.cpp:257-276`` (``u32 777, i32 stype, i32 n``).
* Bond      -- ``src/Bond.cpp:500-607`` (``u32 666, i32 ver``, ...).
* UniTensor -- ``src/UniTensor.cpp:68-167``; block payload ``src/BlockUniTensor.cpp:1618-1662``;
  dense payload is one Tensor (``src/DenseUniTensor.cpp:1314-1316``).

Everything is decoded into plain numpy arrays; nothing in a file is executed.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

# dtype ids follow the reference's Type_list order (include/Type.hpp:177-179).
DTYPE_VOID, DTYPE_C128, DTYPE_C64, DTYPE_F64, DTYPE_F32 = 0, 1, 2, 3, 4
DTYPE_I64, DTYPE_U64, DTYPE_I32, DTYPE_U32, DTYPE_I16, DTYPE_U16, DTYPE_BOOL = 5, 6, 7, 8, 9, 10, 11

NP_DTYPE = {
    DTYPE_C128: np.complex128,
    DTYPE_C64: np.complex64,
    DTYPE_F64: np.float64,
    DTYPE_F32: np.float32,
    DTYPE_I64: np.int64,
    DTYPE_U64: np.uint64,
    DTYPE_I32: np.int32,
    DTYPE_U32: np.uint32,
    DTYPE_I16: np.int16,
    DTYPE_U16: np.uint16,
    DTYPE_BOOL: np.bool_,
}

UTEN_DENSE, UTEN_BLOCK, UTEN_BLOCK_FERMIONIC = 0, 2, 3
BD_KET, BD_BRA, BD_REG = -1, 1, 0
SYM_U1, SYM_ZN = -1, 0


class _Cursor:
    def __init__(self, buf: bytes):
        self.buf = buf
        self.pos = 0

    def take(self, fmt: str):
        size = struct.calcsize("<" + fmt)
        if self.pos + size > len(self.buf):
            raise ValueError("truncated Cytnx file")
        vals = struct.unpack_from("<" + fmt, self.buf, self.pos)
        self.pos += size
        return vals if len(vals) > 1 else vals[0]

    def raw(self, nbytes: int) -> bytes:
        if self.pos + nbytes > len(self.buf):
            raise ValueError("truncated Cytnx file")
        out = self.buf[self.pos:self.pos + nbytes]
        self.pos += nbytes
        return out


@dataclass
class SavedTensor:
    """A Tensor as stored: logical ``shape`` plus the mapper/invmapper of a lazy permute.

    ``data`` holds the storage in memory order.  ``logical()`` applies the stored permutation
    so the result is the row-major array the Tensor represents.
    """
    shape: List[int]
    mapper: List[int]
    invmapper: List[int]
    contiguous: bool
    dtype: int
    data: np.ndarray

    def logical(self) -> np.ndarray:
        if not self.shape:
            return self.data.reshape(())
        # memory shape is shape[invmapper[i]] (include/backend/Tensor_impl.hpp:266-269)
        mem_shape = [self.shape[self.invmapper[i]] for i in range(len(self.shape))]
        return np.transpose(self.data.reshape(mem_shape), self.mapper)


@dataclass
class SavedSymmetry:
    stype: int
    n: int


@dataclass
class SavedBond:
    btype: int
    qnums: List[List[int]]
    degs: List[int]
    syms: List[SavedSymmetry]
    dim: int


@dataclass
class SavedUniTensor:
    uten_type: int
    is_braket_form: bool
    is_tag: bool
    is_diag: bool
    rowrank: int
    name: str
    labels: List[str]
    bonds: List[SavedBond]
    blocks: List[SavedTensor] = field(default_factory=list)
    itoi: List[List[int]] = field(default_factory=list)
    syms_cache: Optional[List[SavedSymmetry]] = None


def _read_storage(c: _Cursor) -> tuple:
    magic = c.take("I")
    if magic != 999:
        raise ValueError("not a Cytnx Storage (magic %d)" % magic)
    size = c.take("Q")
    dtype = c.take("I")
    _device = c.take("i")
    npdt = np.dtype(NP_DTYPE[dtype]).newbyteorder("<")
    data = np.frombuffer(c.raw(size * npdt.itemsize), dtype=npdt).copy()
    return dtype, data


def _read_tensor(c: _Cursor) -> SavedTensor:
    magic = c.take("I")
    if magic == 889:
        version = c.take("I")
        if version != 1:
            raise ValueError("unsupported Tensor file version %d" % version)
    elif magic != 888:
        raise ValueError("not a Cytnx Tensor (magic %d)" % magic)
    rank = c.take("Q")
    conti = c.take("Q")
    shape = list(struct.unpack_from("<%dQ" % rank, c.raw(8 * rank))) if rank else []
    mapper = list(struct.unpack_from("<%dQ" % rank, c.raw(8 * rank))) if rank else []
    invmapper = list(struct.unpack_from("<%dQ" % rank, c.raw(8 * rank))) if rank else []
    dtype, data = _read_storage(c)
    return SavedTensor(shape, mapper, invmapper, bool(conti), dtype, data)


def _read_symmetry(c: _Cursor) -> SavedSymmetry:
    magic = c.take("I")
    if magic != 777:
        raise ValueError("not a Cytnx Symmetry (magic %d)" % magic)
    return SavedSymmetry(c.take("i"), c.take("i"))


def _read_bond(c: _Cursor) -> SavedBond:
    magic = c.take("I")
    if magic != 666:
        raise ValueError("not a Cytnx Bond (magic %d)" % magic)
    ver = c.take("i")
    n = c.take("Q")
    btype = c.take("i")
    nsym = c.take("Q")
    qnums: List[List[int]] = []
    degs: List[int] = []
    if ver == 0:
        dim = n
        if nsym:
            for _ in range(dim):
                qnums.append(list(struct.unpack_from("<%dq" % nsym, c.raw(8 * nsym))))
    else:
        for _ in range(n):
            qnums.append(list(struct.unpack_from("<%dq" % nsym, c.raw(8 * nsym))))
        degs = list(struct.unpack_from("<%dQ" % n, c.raw(8 * n))) if n else []
        dim = int(sum(degs))
    syms = [_read_symmetry(c) for _ in range(nsym)]
    return SavedBond(btype, qnums, degs, syms, dim)


def read_tensor(path: str) -> SavedTensor:
    with open(path, "rb") as f:
        c = _Cursor(f.read())
    return _read_tensor(c)


def read_unitensor(path: str) -> SavedUniTensor:
    with open(path, "rb") as f:
        c = _Cursor(f.read())
    magic = c.take("I")
    version = 0
    if magic == 556:
        version = c.take("I")
        if version != 1:
            raise ValueError("unsupported UniTensor file version %d" % version)
    elif magic != 555:
        raise ValueError("not a Cytnx UniTensor (magic %d)" % magic)
    uten_type = c.take("i")
    braket, is_tag, is_diag = c.take("B"), c.take("B"), c.take("B")
    rowrank = c.take("q")
    len_name = c.take("I")
    name = c.raw(len_name).decode() if len_name else ""
    rank = c.take("Q")
    lens = [c.take("Q") for _ in range(rank)]
    labels = [c.raw(n).decode() for n in lens]
    bonds = [_read_bond(c) for _ in range(rank)]
    ut = SavedUniTensor(uten_type, bool(braket), bool(is_tag), bool(is_diag), rowrank, name,
                        labels, bonds)
    if uten_type == UTEN_DENSE:
        ut.blocks = [_read_tensor(c)]
    elif uten_type in (UTEN_BLOCK, UTEN_BLOCK_FERMIONIC):
        if rank == 0 and version > 0:
            nsym = c.take("Q")
            ut.syms_cache = [_read_symmetry(c) for _ in range(nsym)]
        nblocks = c.take("Q")
        for _ in range(nblocks):
            if rank:
                ut.itoi.append(list(struct.unpack_from("<%dQ" % rank, c.raw(8 * rank))))
            else:
                ut.itoi.append([])
        ut.blocks = [_read_tensor(c) for _ in range(nblocks)]
    else:
        raise ValueError("unsupported UniTensor type %d" % uten_type)
    if c.pos != len(c.buf):
        raise ValueError("trailing bytes in Cytnx file")
    return ut
