"""Dense Tensor with the reference's lazy-permute semantics.

Mirrors ``cytnx::Tensor`` / ``Tensor_impl`` (include/Tensor.hpp:718-824,
include/backend/Tensor_impl.hpp:109-325, src/backend/Tensor_impl.cpp:38-143):

* ``permute_`` / ``permute`` only compose the mapper/invmapper metadata; no data moves;
* ``contiguous_`` / ``contiguous`` move the data once, through ``cxa_permute`` on the GPU;
* ``reshape_`` forces ``contiguous_`` first.

Storage is a flat PyTorch buffer (device memory on a HIP device).  Physical moves on a CPU
tensor are refused: this package implements the Device.cuda branch only.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np
import torch

from . import _lib
from .device import CPU, NP_OF, TORCH_OF, dtype_id, stream_handle, torch_device


def _flatten_args(args) -> List[int]:
    if len(args) == 1 and isinstance(args[0], (list, tuple)):
        return [int(x) for x in args[0]]
    return [int(x) for x in args]


class Tensor:
    __slots__ = ("_storage", "_shape", "_mapper", "_invmapper", "_contiguous")

    def __init__(self, shape: Sequence[int] = (), dtype: int = 3, device: int = CPU,
                 init_zero: bool = True):
        shape = [int(s) for s in shape]
        n = int(np.prod(shape)) if shape else 1
        alloc = torch.zeros if init_zero else torch.empty
        self._storage = alloc(n, dtype=TORCH_OF[dtype], device=torch_device(device))
        self._shape = shape
        self._mapper = list(range(len(shape)))
        self._invmapper = list(range(len(shape)))
        self._contiguous = True

    # ---------------------------------------------------------------- construction helpers
    @classmethod
    def from_storage(cls, flat: torch.Tensor, shape: Sequence[int]) -> "Tensor":
        t = cls.__new__(cls)
        # the kernels read raw memory: a lazily conjugated or negated torch view (torch's SVD
        # returns Vh as one) is materialised here
        flat = flat.resolve_conj().resolve_neg()
        t._storage = flat.reshape(-1)
        t._shape = [int(s) for s in shape]
        t._mapper = list(range(len(t._shape)))
        t._invmapper = list(range(len(t._shape)))
        t._contiguous = True
        if int(np.prod(t._shape)) != t._storage.numel() and t._shape:
            raise ValueError("storage size does not match shape")
        return t

    @classmethod
    def from_numpy(cls, a: np.ndarray, device: int = CPU) -> "Tensor":
        a = np.ascontiguousarray(a)
        flat = torch.from_numpy(a.reshape(-1).copy()).to(torch_device(device))
        return cls.from_storage(flat, a.shape)

    # ---------------------------------------------------------------- metadata
    def shape(self) -> List[int]:
        return list(self._shape)

    def rank(self) -> int:
        return len(self._shape)

    def dtype(self) -> int:
        return dtype_id(self._storage.dtype)

    def device(self) -> int:
        return CPU if self._storage.device.type == "cpu" else self._storage.device.index

    def is_contiguous(self) -> bool:
        return self._contiguous

    def mapper(self) -> List[int]:
        return list(self._mapper)

    def invmapper(self) -> List[int]:
        return list(self._invmapper)

    def storage(self) -> torch.Tensor:
        return self._storage

    def same_data(self, other: "Tensor") -> bool:
        return self._storage.data_ptr() == other._storage.data_ptr()

    # ---------------------------------------------------------------- lazy permute
    def _check_perm(self, rnks: List[int]) -> None:
        if len(rnks) != len(self._shape):
            raise ValueError("reshape a tensor with a specify shape that does not match with the "
                             "shape of the incident tensor.")
        if len(set(rnks)) != len(rnks):
            raise ValueError("tensor permute with duplicated index.")
        if any(r < 0 or r >= len(rnks) for r in rnks):
            raise ValueError("reshape a tensor with invalid rank index.")

    def permute_(self, *args) -> "Tensor":
        """Tensor_impl::permute_ (src/backend/Tensor_impl.cpp:88-143)."""
        rnks = _flatten_args(args)
        self._check_perm(rnks)
        new_map = [self._mapper[r] for r in rnks]
        new_shape = [self._shape[r] for r in rnks]
        for i, r in enumerate(rnks):
            self._invmapper[self._mapper[r]] = i
        self._mapper = new_map
        self._shape = new_shape
        self._contiguous = all(m == i for i, m in enumerate(self._mapper))
        return self

    def permute(self, *args) -> "Tensor":
        """Tensor_impl::permute (src/backend/Tensor_impl.cpp:38-86): shares the storage."""
        out = Tensor.__new__(Tensor)
        out._storage = self._storage
        out._shape = list(self._shape)
        out._mapper = list(self._mapper)
        out._invmapper = list(self._invmapper)
        out._contiguous = self._contiguous
        return out.permute_(*args)

    # ---------------------------------------------------------------- physical moves
    def _moved_storage(self) -> torch.Tensor:
        if self._storage.numel() == 0:
            return self._storage
        if self._storage.device.type == "cpu":
            raise NotImplementedError(
                "cytnx_amd implements the Device.cuda branch; move the tensor to a GPU first")
        oldshape = [self._shape[self._invmapper[i]] for i in range(len(self._shape))]
        dst = torch.empty_like(self._storage)
        L = _lib.lib()
        _lib.check(L.cxa_permute(self.dtype(), len(oldshape), _lib.u64_array(oldshape),
                                 _lib.u64_array(self._mapper), self._storage.data_ptr(),
                                 dst.data_ptr(), stream_handle(self._storage.device.index)))
        return dst

    def contiguous_(self) -> "Tensor":
        """Tensor_impl::contiguous_ (include/backend/Tensor_impl.hpp:256-279)."""
        if not self._contiguous:
            self._storage = self._moved_storage()
            self._mapper = list(range(len(self._shape)))
            self._invmapper = list(range(len(self._shape)))
            self._contiguous = True
        return self

    def contiguous(self) -> "Tensor":
        """Tensor_impl::contiguous (include/backend/Tensor_impl.hpp:222-254): self if already
        contiguous, otherwise a new tensor with moved data."""
        if self._contiguous:
            return self
        return Tensor.from_storage(self._moved_storage(), self._shape)

    def reshape_(self, *args) -> "Tensor":
        """Tensor_impl::reshape_ (include/backend/Tensor_impl.hpp:281-325)."""
        new_shape = _flatten_args(args)
        n = self._storage.numel()
        known, und = 1, -1
        for i, s in enumerate(new_shape):
            if s < 0:
                if s != -1:
                    raise ValueError("reshape can only have dimension > 0 and one undetermine "
                                     "rank specify as -1")
                if und >= 0:
                    raise ValueError("reshape can have only one dimension specified as -1")
                und = i
            else:
                known *= s
        if und >= 0:
            if known == 0 or n % known:
                raise ValueError("unmatch size when reshape with undetermine dimension")
            new_shape[und] = n // known
        elif known != n:
            raise ValueError("new shape does not match the number of elements.")
        self.contiguous_()
        self._shape = new_shape
        self._mapper = list(range(len(new_shape)))
        self._invmapper = list(range(len(new_shape)))
        return self

    def reshape(self, *args) -> "Tensor":
        out = self.contiguous() if not self._contiguous else self.clone_meta()
        return out.reshape_(*args)

    def clone_meta(self) -> "Tensor":
        out = Tensor.__new__(Tensor)
        out._storage = self._storage
        out._shape = list(self._shape)
        out._mapper = list(self._mapper)
        out._invmapper = list(self._invmapper)
        out._contiguous = self._contiguous
        return out

    def clone(self) -> "Tensor":
        out = self.clone_meta()
        out._storage = self._storage.clone()
        return out

    def to(self, device: int) -> "Tensor":
        out = self.clone_meta()
        out._storage = self._storage.to(torch_device(device))
        return out

    # ---------------------------------------------------------------- host views
    def numpy(self) -> np.ndarray:
        """Logical row-major array (host copy)."""
        mem = self._storage.detach().cpu().numpy()
        if not self._shape:
            return mem.reshape(())
        mem_shape = [self._shape[self._invmapper[i]] for i in range(len(self._shape))]
        return np.ascontiguousarray(np.transpose(mem.reshape(mem_shape), self._mapper))

    def torch(self) -> torch.Tensor:
        """The contiguous logical data as a torch tensor (moves data if needed)."""
        t = self.contiguous()
        return t._storage.reshape(t._shape)

    # ---------------------------------------------------------------- files
    def Save(self, fname: str) -> None:
        """Tensor::Save (src/Tensor.cpp:452-470): '.cytn' is appended to a name without an
        extension; the tensor is written contiguous."""
        import os
        from . import fileio
        if not os.path.splitext(fname)[1]:
            fname = fname + ".cytn"
        with open(fname, "wb") as f:
            f.write(fileio.tensor_bytes(self.numpy()))

    @classmethod
    def Load(cls, fname: str, device: int = CPU) -> "Tensor":
        """Tensor::Load: the stored (possibly lazily permuted) tensor, contiguous."""
        from . import fileio
        return cls.from_numpy(fileio.read_tensor(fname).logical(), device)

    def __repr__(self) -> str:
        return "Tensor(shape=%s, dtype=%d, device=%d, contiguous=%s)" % (
            self._shape, self.dtype(), self.device(), self._contiguous)


def from_numpy(a: np.ndarray, device: int = CPU) -> Tensor:
    return Tensor.from_numpy(a, device)


__all__ = ["Tensor", "from_numpy", "NP_OF"]
