"""cxa_tensordot_noalloc, the tensordot entry point that allocates nothing (its permuted operand
copies go to the library's static device scratch), through the C ABI: equal bit for bit to
cxa_tensordot with a caller workspace and within 1e-12 of the oracle's tensordot, on two calls
in a row with the second on smaller operands (the scratch keeps the first call's contents)."""
import ctypes

import numpy as np
import pytest
import torch

import pyoracle as po
from cytnx_amd import _lib

pytestmark = pytest.mark.gpu

DT = {np.float64: 3, np.complex128: 1}


def _rand(rng, shape, dtype):
    x = rng.standard_normal(shape)
    if dtype == np.complex128:
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(dtype)


def _tensordot(noalloc, L, R, il, ir):
    """Run one entry point on device copies of L and R; the output as a numpy array."""
    dt = DT[L.dtype.type]
    dL, dR = torch.from_numpy(L).cuda(), torch.from_numpy(R).cuda()
    fl = [i for i in range(L.ndim) if i not in il]
    fr = [i for i in range(R.ndim) if i not in ir]
    out = torch.empty([L.shape[i] for i in fl] + [R.shape[i] for i in fr], dtype=dL.dtype,
                      device="cuda")
    sl, sr = _lib.u64_array(L.shape), _lib.u64_array(R.shape)
    cl, cr = _lib.u64_array(il), _lib.u64_array(ir)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    lib = _lib.lib()
    if noalloc:
        _lib.check(lib.cxa_tensordot_noalloc(dt, out.data_ptr(), dL.data_ptr(), L.ndim, sl,
                                             dR.data_ptr(), R.ndim, sr, len(il), cl, cr, stream))
    else:
        ws = ctypes.c_int64(0)
        _lib.check(lib.cxa_tensordot_workspace(dt, L.ndim, sl, R.ndim, sr, len(il), cl, cr,
                                               ctypes.byref(ws)))
        assert ws.value > 0  # these cases copy both operands
        wbuf = torch.empty(ws.value, dtype=torch.uint8, device="cuda")
        _lib.check(lib.cxa_tensordot(dt, out.data_ptr(), dL.data_ptr(), L.ndim, sl,
                                     dR.data_ptr(), R.ndim, sr, len(il), cl, cr,
                                     wbuf.data_ptr(), stream))
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_tensordot_noalloc_matches_tensordot_twice(dtype):
    rng = np.random.default_rng(71)
    # both operands need a permuted copy: L's free axes (1, 3) are not adjacent, and R's
    # contracted axes (0, 2) are not one strided index
    cases = [((12, 10, 14, 16), (14, 9, 12), [2, 0], [0, 2]),
             ((3, 4, 5, 6), (5, 2, 3), [2, 0], [0, 2])]
    for sl, sr, il, ir in cases:
        print("tensordot_noalloc", np.dtype(dtype).name, sl, sr, il, ir, flush=True)
        L, R = _rand(rng, sl, dtype), _rand(rng, sr, dtype)
        want = _tensordot(False, L, R, il, ir)
        got = _tensordot(True, L, R, il, ir)
        assert np.array_equal(got, want)
        ref = po.tensordot(L, R, il, ir)
        assert np.linalg.norm(got - ref) <= 1e-12 * np.linalg.norm(ref)
