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


# ------------------------------------------------------------------------------------------
# Writer: the inverse of the reader, following the reference's _Save chain (UniTensor::_Save,
# src/UniTensor.cpp:68-110; BlockUniTensor::_save_dispatch, src/BlockUniTensor.cpp:1618-1635;
# Bond::_Save, src/Bond.cpp:500-548; Tensor::_Save, src/Tensor.cpp:472-493; Storage::_Save,
# src/backend/Storage.cpp:161-193).  Tensors are written contiguous (identity mapper) and every
# Storage as a host storage (device -1), so the file loads on any build.
# ------------------------------------------------------------------------------------------
DTYPE_OF_NP = {np.dtype(v): k for k, v in NP_DTYPE.items()}


def _w(out: bytearray, fmt: str, *vals) -> None:
    out += struct.pack("<" + fmt, *vals)


def _write_storage(out: bytearray, data: np.ndarray) -> None:
    flat = np.ascontiguousarray(data).reshape(-1)
    dtype = DTYPE_OF_NP[flat.dtype]
    _w(out, "IQIi", 999, flat.size, dtype, -1)
    out += flat.astype(flat.dtype.newbyteorder("<"), copy=False).tobytes()


def _write_tensor(out: bytearray, arr: np.ndarray, legacy: bool = False) -> None:
    arr = np.ascontiguousarray(arr)
    rank = arr.ndim
    if legacy:
        _w(out, "IQQ", 888, rank, 1)
    else:
        _w(out, "IIQQ", 889, 1, rank, 1)
    if rank:
        ident = list(range(rank))
        _w(out, "%dQ" % rank, *[int(s) for s in arr.shape])
        _w(out, "%dQ" % rank, *ident)
        _w(out, "%dQ" % rank, *ident)
    _write_storage(out, arr)


def _write_symmetry(out: bytearray, s: SavedSymmetry) -> None:
    _w(out, "Iii", 777, s.stype, s.n)


def _write_bond(out: bytearray, b: SavedBond) -> None:
    _w(out, "I", 666)
    ver = 1 if b.degs else 0
    nsym = len(b.syms)
    _w(out, "iQiQ", ver, len(b.qnums) if ver else b.dim, b.btype, nsym)
    for q in b.qnums:
        _w(out, "%dq" % nsym, *q)
    if ver:
        _w(out, "%dQ" % len(b.degs), *b.degs)
    for s in b.syms:
        _write_symmetry(out, s)


def unitensor_bytes(ut: SavedUniTensor, legacy: bool = False) -> bytes:
    """The file image of `ut` (blocks given as logical row-major arrays in ut.blocks).

    The default is the versioned format the reference writes today (UniTensor magic 556 and
    Tensor magic 889, each followed by format version 1).  `legacy=True` writes the unversioned
    format (magics 555 / 888, no version words) that older Cytnx releases write and read; the
    reference's loader accepts both (src/UniTensor.cpp:112-125, src/Tensor.cpp:516-530)."""
    out = bytearray()
    if legacy:
        _w(out, "I", 555)
    else:
        _w(out, "II", 556, 1)
    _w(out, "i???q", ut.uten_type, ut.is_braket_form, ut.is_tag, ut.is_diag, ut.rowrank)
    name = ut.name.encode()
    _w(out, "I", len(name))
    out += name
    labels = [l.encode() for l in ut.labels]
    _w(out, "Q", len(labels))
    for l in labels:
        _w(out, "Q", len(l))
    for l in labels:
        out += l
    for b in ut.bonds:
        _write_bond(out, b)
    if ut.uten_type == UTEN_DENSE:
        _write_tensor(out, ut.blocks[0], legacy)
    elif ut.uten_type == UTEN_BLOCK:
        if not ut.labels and not legacy:
            syms = ut.syms_cache or []
            _w(out, "Q", len(syms))
            for s in syms:
                _write_symmetry(out, s)
        _w(out, "Q", len(ut.blocks))
        for t in ut.itoi:
            if t:
                _w(out, "%dQ" % len(t), *t)
        for blk in ut.blocks:
            _write_tensor(out, blk, legacy)
    else:
        raise ValueError("only dense and block UniTensors are written")
    return bytes(out)


def write_unitensor(path: str, ut: SavedUniTensor, legacy: bool = False) -> None:
    with open(path, "wb") as f:
        f.write(unitensor_bytes(ut, legacy))


def tensor_bytes(arr: np.ndarray, legacy: bool = False) -> bytes:
    """The .cytn image of a logical row-major array (Tensor::_Save, src/Tensor.cpp:472-493)."""
    out = bytearray()
    _write_tensor(out, arr, legacy)
    return bytes(out)
