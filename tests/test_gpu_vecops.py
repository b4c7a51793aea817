"""The vector drop-ins cxa_dot (Vectordotfunc_oii) and cxa_nrm2 (Normfunc_oii) on F64, F32,
C128 and C64, through linalg.Vectordot / linalg.norm / UniTensor.Norm, against numpy in
float64 / complex128 (products summed in FP64 on the device, so F32 / C64 results are the
rounded FP64 sums); edge cases n = 0, 1 and lengths that leave partial workgroups."""
import numpy as np
import pytest
import torch

from cytnx_amd import _lib, heisenberg, linalg
from cytnx_amd.device import stream_handle
from cytnx_amd.tensor import Tensor

pytestmark = pytest.mark.gpu

DT = {np.float64: 3, np.float32: 4, np.complex128: 1, np.complex64: 2}


def _rand(rng, n, dtype):
    x = rng.standard_normal(n)
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * rng.standard_normal(n)
    return x.astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
@pytest.mark.parametrize("n", [1, 255, 257, 131072 + 17, 231768])
def test_vectordot_and_norm(dtype, n):
    rng = np.random.default_rng(n)
    x, y = _rand(rng, n, dtype), _rand(rng, n, dtype)
    wide = np.complex128 if np.issubdtype(dtype, np.complexfloating) else np.float64
    X, Y = x.astype(wide), y.astype(wide)
    tol = 1e-13 if dtype in (np.float64, np.complex128) else 1e-7
    for conj in (False, True):
        got = linalg.Vectordot(Tensor.from_numpy(x, 0), Tensor.from_numpy(y, 0), conj).numpy()
        ref = np.sum((np.conj(X) if conj else X) * Y)
        assert got.dtype == dtype
        assert abs(got[0] - ref) <= tol * np.sum(np.abs(X) * np.abs(Y))
    nrm = linalg.norm(Tensor.from_numpy(x, 0))
    assert abs(nrm - np.linalg.norm(X)) <= tol * np.linalg.norm(X)
    with pytest.warns(DeprecationWarning):
        N = linalg.Norm(Tensor.from_numpy(x, 0)).numpy()
    assert N.dtype == (np.float64 if dtype in (np.float64, np.complex128) else np.float32)


def test_dot_is_deterministic_and_raw_abi():
    rng = np.random.default_rng(3)
    x = torch.from_numpy(rng.standard_normal(1 << 20)).cuda()
    ws = linalg.vec_workspace(0)
    out = torch.empty(2, dtype=torch.float64, device="cuda:0")
    L = _lib.lib()
    _lib.check(L.cxa_dot(3, x.numel(), x.data_ptr(), x.data_ptr(), 0, out.data_ptr(),
                         ws.data_ptr(), stream_handle(0)))
    _lib.check(L.cxa_nrm2(3, x.numel(), x.data_ptr(), out.data_ptr() + 8, ws.data_ptr(),
                          stream_handle(0)))
    a = out.cpu().numpy().copy()
    for _ in range(3):
        _lib.check(L.cxa_dot(3, x.numel(), x.data_ptr(), x.data_ptr(), 0, out.data_ptr(),
                             ws.data_ptr(), stream_handle(0)))
        assert out.cpu().numpy()[0] == a[0]
    assert abs(a[1] - np.sqrt(a[0])) <= 1e-15 * a[1]
    # n = 0: the dot is 0 and the norm 0
    _lib.check(L.cxa_dot(3, 0, x.data_ptr(), x.data_ptr(), 0, out.data_ptr(), ws.data_ptr(),
                         stream_handle(0)))
    assert out.cpu().numpy()[0] == 0.0
    with pytest.raises(_lib.CxaError):
        _lib.check(L.cxa_dot(5, 4, x.data_ptr(), x.data_ptr(), 0, out.data_ptr(),
                             ws.data_ptr(), stream_handle(0)))


def test_unitensor_norm_through_nrm2():
    psi = heisenberg.two_site_problem(64, 2, True)[0].to(0)
    ref = np.linalg.norm(psi._storage.cpu().numpy())
    assert abs(psi.Norm() - ref) <= 1e-14 * ref
