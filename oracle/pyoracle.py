"""ctypes front-end to the CPU oracle (oracle/oracle.c) plus a small block-tensor model.

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the CPU-baseline leg
of bench.py, never by the product package.

``OBlockTensor`` models a U(1) BlockUniTensor just far enough to drive the oracle: bonds
(qnums, degeneracies, KET/BRA), labels, rowrank, the block list in the reference's
enumeration order and the block data as numpy arrays.
"""
from __future__ import annotations

import ctypes
import itertools
import os
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

F64, C128 = 3, 1
KET, BRA = -1, 1


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(path)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        vp = ctypes.c_void_p
        L.orc_permute.argtypes = [ctypes.c_int, ctypes.c_int, u64p, u64p, vp, vp]
        L.orc_matmul.argtypes = [ctypes.c_int, vp, vp, vp, ctypes.c_int64, ctypes.c_int64,
                                 ctypes.c_int64]
        L.orc_tensordot.argtypes = [ctypes.c_int, ctypes.c_int, u64p, vp, ctypes.c_int, u64p, vp,
                                    ctypes.c_int, u64p, u64p, vp]
        L.orc_block_contract.argtypes = [ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, u64p, u64p, vp,
                                         ctypes.c_int, ctypes.c_int, u64p, u64p, vp,
                                         ctypes.c_int, u64p, u64p,
                                         ctypes.c_int, u64p, u64p, vp]
        L.orc_block_fullcontract.argtypes = [ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, u64p, u64p, vp,
                                             ctypes.c_int, u64p, u64p, vp, u64p, u64p, vp]
        L.orc_tridiag_eigh.argtypes = [ctypes.c_int, vp, vp, vp, vp]
        L.orc_set_dgemm.argtypes = [vp]
        L.orc_lanczos_gnd.argtypes = [vp, vp, ctypes.c_int64, vp, ctypes.c_double, ctypes.c_int,
                                      vp, vp]
        _LIB = L
    return _LIB


def _u64(seq) -> ctypes.Array:
    seq = [int(x) for x in seq]
    return (ctypes.c_uint64 * max(1, len(seq)))(*seq)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _dtype_id(a: np.ndarray) -> int:
    if a.dtype == np.float64:
        return F64
    if a.dtype == np.complex128:
        return C128
    raise TypeError("oracle supports float64 and complex128, got %s" % a.dtype)


# --------------------------------------------------------------------------------------------
# dense primitives
# --------------------------------------------------------------------------------------------
def permute(a: np.ndarray, mapper: Sequence[int]) -> np.ndarray:
    """Physical permute of a C-contiguous array: result[j...] axis j = source axis mapper[j]."""
    a = np.ascontiguousarray(a)
    out = np.empty([a.shape[m] for m in mapper], dtype=a.dtype)
    if a.size:
        rc = lib().orc_permute(a.dtype.itemsize, a.ndim, _u64(a.shape), _u64(mapper), _ptr(a),
                               _ptr(out))
        assert rc == 0
    return out


def matmul(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    A = np.ascontiguousarray(A)
    B = np.ascontiguousarray(B.astype(A.dtype))
    out = np.empty((A.shape[0], B.shape[1]), dtype=A.dtype)
    assert lib().orc_matmul(_dtype_id(A), _ptr(out), _ptr(A), _ptr(B), A.shape[0], A.shape[1],
                            B.shape[1]) == 0
    return out


def tensordot(L: np.ndarray, R: np.ndarray, idxl: Sequence[int], idxr: Sequence[int]):
    L = np.ascontiguousarray(L)
    R = np.ascontiguousarray(R.astype(L.dtype))
    fl = [i for i in range(L.ndim) if i not in idxl]
    fr = [i for i in range(R.ndim) if i not in idxr]
    out = np.empty([L.shape[i] for i in fl] + [R.shape[i] for i in fr], dtype=L.dtype)
    rc = lib().orc_tensordot(_dtype_id(L), L.ndim, _u64(L.shape), _ptr(L), R.ndim, _u64(R.shape),
                             _ptr(R), len(idxl), _u64(idxl), _u64(idxr), _ptr(out))
    assert rc == 0, rc
    return out


def tridiag_eigh(d: np.ndarray, e: np.ndarray):
    d = np.ascontiguousarray(d, dtype=np.float64)
    e = np.ascontiguousarray(e, dtype=np.float64)
    n = d.size
    w = np.empty(n)
    Z = np.empty((n, n))
    rc = lib().orc_tridiag_eigh(n, _ptr(d), _ptr(e) if e.size else None, _ptr(w), _ptr(Z))
    return w, Z, rc


MATVEC_T = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def lanczos_gnd(matvec, Tin: np.ndarray, CvgCrit: float = 1e-14, Maxiter: int = 10000):
    """Ground state by the restated Lanczos_Gnd_Ut; matvec maps a float64 vector to H @ v."""
    n = Tin.size
    Tin = np.ascontiguousarray(Tin, dtype=np.float64).ravel()

    def cb(xp, yp, _ctx):
        x = np.ctypeslib.as_array(xp, shape=(n,))
        y = np.ctypeslib.as_array(yp, shape=(n,))
        y[:] = matvec(x.copy())

    cfn = MATVEC_T(cb)
    E = ctypes.c_double(0.0)
    v = np.empty(n)
    na = lib().orc_lanczos_gnd(ctypes.cast(cfn, ctypes.c_void_p), None, n, _ptr(Tin),
                               CvgCrit, Maxiter, ctypes.addressof(E), _ptr(v))
    if na < 0:
        raise ValueError("oracle Lanczos rejected its arguments (%d)" % na)
    return E.value, v, na


# --------------------------------------------------------------------------------------------
# block tensors
# --------------------------------------------------------------------------------------------
class OBond:
    def __init__(self, btype: int, qnums: Sequence[int], degs: Sequence[int]):
        self.type = int(btype)
        self.qnums = [int(q) for q in qnums]
        self.degs = [int(d) for d in degs]

    def redirect(self) -> "OBond":
        return OBond(-self.type, self.qnums, self.degs)

    @property
    def dim(self) -> int:
        return int(sum(self.degs))

    def offsets(self) -> List[int]:
        return list(np.concatenate([[0], np.cumsum(self.degs)[:-1]]).astype(int))


def enumerate_blocks(bonds: Sequence[OBond]) -> List[List[int]]:
    """Zero-flux qindex tuples in odometer (lexicographic) order, by brute force.

    Follows BlockUniTensor::Init (src/BlockUniTensor.cpp:190-230) with the U(1) flux rule of
    include/UniTensor.hpp:1139-1166 (KET adds +q, BRA adds -q; src/Symmetry.cpp:33-50).
    """
    out = []
    for loc in itertools.product(*[range(len(b.qnums)) for b in bonds]):
        flux = 0
        for b, i in zip(bonds, loc):
            flux += b.qnums[i] if b.type == KET else -b.qnums[i]
        if flux == 0:
            out.append(list(loc))
    return out


class OBlockTensor:
    def __init__(self, bonds: Sequence[OBond], labels: Sequence[str], rowrank: int,
                 dtype=np.float64, itoi: Optional[List[List[int]]] = None,
                 blocks: Optional[List[np.ndarray]] = None):
        self.bonds = list(bonds)
        self.labels = [str(x) for x in labels]
        self.rowrank = int(rowrank)
        self.itoi = [list(t) for t in (itoi if itoi is not None else enumerate_blocks(bonds))]
        if blocks is None:
            blocks = [np.zeros([b.degs[i] for b, i in zip(self.bonds, t)], dtype=dtype)
                      for t in self.itoi]
        self.blocks = [np.ascontiguousarray(b) for b in blocks]

    @property
    def rank(self) -> int:
        return len(self.bonds)

    @property
    def dtype(self):
        return self.blocks[0].dtype if self.blocks else np.dtype(np.float64)

    def shape(self) -> List[int]:
        return [b.dim for b in self.bonds]

    def permute(self, mapper: Sequence[int], rowrank: Optional[int] = None) -> "OBlockTensor":
        """BlockUniTensor::permute (src/BlockUniTensor.cpp:565-633): permute bonds, labels,
        every itoi tuple and every block (block order unchanged)."""
        return OBlockTensor([self.bonds[m] for m in mapper], [self.labels[m] for m in mapper],
                            self.rowrank if rowrank is None else rowrank, self.dtype,
                            [[t[m] for m in mapper] for t in self.itoi],
                            [permute(b, mapper) for b in self.blocks])

    def to_dense(self) -> np.ndarray:
        out = np.zeros(self.shape(), dtype=self.dtype)
        offs = [b.offsets() for b in self.bonds]
        for t, blk in zip(self.itoi, self.blocks):
            sl = tuple(slice(offs[a][q], offs[a][q] + blk.shape[a]) for a, q in enumerate(t))
            out[sl] = blk
        return out

    def exists_mask(self) -> np.ndarray:
        out = np.zeros(self.shape(), dtype=bool)
        offs = [b.offsets() for b in self.bonds]
        for t, blk in zip(self.itoi, self.blocks):
            sl = tuple(slice(offs[a][q], offs[a][q] + blk.shape[a]) for a, q in enumerate(t))
            out[sl] = True
        return out

    def norm(self) -> float:
        return float(np.sqrt(sum(float(np.vdot(b, b).real) for b in self.blocks)))

    def flat(self) -> np.ndarray:
        return np.concatenate([b.ravel() for b in self.blocks]) if self.blocks else np.zeros(0)

    def set_flat(self, v: np.ndarray) -> None:
        o = 0
        for i, b in enumerate(self.blocks):
            self.blocks[i] = np.ascontiguousarray(v[o:o + b.size].reshape(b.shape))
            o += b.size


def common_labels(la: Sequence[str], lb: Sequence[str]):
    """vec_intersect_ (src/utils/vec_intersect.cpp:35-51): common labels sorted
    lexicographically; indices follow that sorted order on both sides."""
    comm = sorted(set(la) & set(lb))
    return comm, [list(la).index(c) for c in comm], [list(lb).index(c) for c in comm]


def _flat_meta(T: OBlockTensor):
    itoi = _u64([q for t in T.itoi for q in t])
    shp = _u64([s for b in T.blocks for s in b.shape])
    ptrs = (ctypes.c_void_p * max(1, len(T.blocks)))(*[_ptr(b) for b in T.blocks])
    return itoi, shp, ptrs


def block_contract(A: OBlockTensor, B: OBlockTensor):
    """UniTensor::contract on two U(1) block tensors (src/BlockUniTensor.cpp:777-1235).

    Returns an OBlockTensor, or a Python scalar when every leg is contracted."""
    comm, c1, c2 = common_labels(A.labels, B.labels)
    if not comm:
        raise NotImplementedError("no-common-label (Kron) case is not part of the oracle")
    dt = np.result_type(A.dtype, B.dtype)
    if dt not in (np.float64, np.complex128):
        raise TypeError(dt)
    A2 = A if A.dtype == dt else OBlockTensor(A.bonds, A.labels, A.rowrank, dt, A.itoi,
                                              [b.astype(dt) for b in A.blocks])
    B2 = B if B.dtype == dt else OBlockTensor(B.bonds, B.labels, B.rowrank, dt, B.itoi,
                                              [b.astype(dt) for b in B.blocks])
    did = F64 if dt == np.float64 else C128
    if len(comm) == A.rank and len(comm) == B.rank:
        out = np.zeros(1, dtype=dt)
        iL, sL, pL = _flat_meta(A2)
        iR, sR, pR = _flat_meta(B2)
        rc = lib().orc_block_fullcontract(did, A.rank, len(A2.blocks), iL, sL, pL,
                                          len(B2.blocks), iR, sR, pR, _u64(c1), _u64(c2),
                                          _ptr(out))
        assert rc == 0
        return out[0]
    nc1 = [i for i in range(A.rank) if i not in c1]
    nc2 = [i for i in range(B.rank) if i not in c2]
    out_bonds = [A.bonds[i] for i in nc1] + [B.bonds[i] for i in nc2]
    out_labels = [A.labels[i] for i in nc1] + [B.labels[i] for i in nc2]
    rr = A.rowrank + B.rowrank - sum(1 for i in c1 if i < A.rowrank) \
        - sum(1 for i in c2 if i < B.rowrank)
    C = OBlockTensor(out_bonds, out_labels, rr, dt)
    iL, sL, pL = _flat_meta(A2)
    iR, sR, pR = _flat_meta(B2)
    iC, sC, pC = _flat_meta(C)
    rc = lib().orc_block_contract(did, A.rank, len(A2.blocks), iL, sL, pL,
                                  B.rank, len(B2.blocks), iR, sR, pR,
                                  len(comm), _u64(c1), _u64(c2),
                                  len(C.blocks), iC, sC, pC)
    assert rc == 0, rc
    return C


def from_saved(ut) -> OBlockTensor:
    """Build an OBlockTensor from a cytnx_amd.fileio.SavedUniTensor (block type)."""
    bonds = [OBond(b.btype, [q[0] for q in b.qnums], b.degs) for b in ut.bonds]
    return OBlockTensor(bonds, ut.labels, ut.rowrank, itoi=ut.itoi,
                        blocks=[np.ascontiguousarray(t.logical()) for t in ut.blocks])


def use_openblas(nthreads: Optional[int] = None) -> str:
    """Route the oracle's F64 GEMM through numpy's bundled OpenBLAS (the reference's default
    CPU BLAS) for timing.  Returns a description of the library, or '' when unavailable."""
    import glob
    here = os.path.dirname(np.__file__)
    cands = sorted(glob.glob(os.path.join(here, "..", "numpy.libs", "libscipy_openblas64_*.so")))
    if not cands:
        return ""
    ob = ctypes.CDLL(cands[0])
    fn = getattr(ob, "scipy_cblas_dgemm64_", None)
    if fn is None:
        return ""
    if nthreads is not None and hasattr(ob, "scipy_openblas_set_num_threads64_"):
        ob.scipy_openblas_set_num_threads64_.argtypes = [ctypes.c_int64]
        ob.scipy_openblas_set_num_threads64_(int(nthreads))
    lib().orc_set_dgemm(ctypes.cast(fn, ctypes.c_void_p))
    return os.path.basename(cands[0])


def use_reference_loops() -> None:
    lib().orc_set_dgemm(None)
