"""k extreme eigenpairs (linalg::Lanczos(Hop, Tin, which, ...), src/linalg/Lanczos.cpp:249-583)
and Lanczos_Exp (src/linalg/Lanczos_Exp.cpp:266-450) on the device drivers cxa_lanczos_eigsh /
cxa_lanczos_exp.

The reference runs the first with ARPACK (absent here), so its Krylov trajectory is not
reproduced; the eigenpairs are pinned by numpy.linalg.eigh of the same dense operators, in the
reference's output order (tests/gpu/linalg_test/Lanczos_test.cpp cases: SA / LA / LM, k = 1,
k = 11 of dim 13, dim 3).  Lanczos_Exp is pinned by the closed forms of the reference's own tests
(tests/linalg_test/Lanczos_Exp_test.cpp: a two-dimensional swap, a three-dimensional operator
with a zero row, a one-dimensional scale, a non-unit start vector) and by scipy.linalg.expm of
random symmetric matrices.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from cytnx_amd.lanczos import LinOp, Lanczos, Lanczos_Exp, lanczos_gnd_native  # noqa: E402
from cytnx_amd.tensor import Tensor  # noqa: E402


class DenseOp(LinOp):
    """y = H x for a dense symmetric H held on the device."""

    def __init__(self, H):
        super().__init__("mv", H.shape[0], 3, 0)
        self.H = torch.as_tensor(H, dtype=torch.float64, device="cuda")

    def matvec(self, v):
        x = v.storage().reshape(-1)
        return Tensor.from_storage(self.H @ x, [x.numel()])


def dev_vec(a):
    t = torch.as_tensor(np.asarray(a, dtype=np.float64), device="cuda")
    return Tensor.from_storage(t, [t.numel()])


def sym(n, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return (A + A.T) / 2


def order(w, which):
    if which == "SA":
        return np.sort(w)
    if which == "LA":
        return np.sort(w)[::-1]
    return w[np.argsort(-np.abs(w), kind="stable")]


@pytest.mark.parametrize("which", ["SA", "LA", "LM"])
@pytest.mark.parametrize("n,k", [(60, 1), (60, 3), (13, 11), (3, 1)])
def test_which_eigenpairs_vs_eigh(which, n, k):
    H = sym(n, 100 + n + k)
    w = np.linalg.eigvalsh(H)
    want = order(w, which)[:k]
    x0 = np.random.default_rng(7).standard_normal(n)
    out = Lanczos(DenseOp(H), dev_vec(x0), which=which, k=k)
    vals = out[0].storage().cpu().numpy()
    scale = max(1.0, np.abs(w).max())
    assert np.allclose(vals, want, atol=1e-10 * scale, rtol=0)
    V = out[1].storage().reshape(k, n).cpu().numpy() if k > 1 else \
        out[1].storage().reshape(1, n).cpu().numpy()
    for i in range(k):
        x = V[i]
        assert abs(np.linalg.norm(x) - 1) < 1e-10
        assert np.linalg.norm(H @ x - vals[i] * x) < 1e-8 * scale
    assert list(out[1].shape()) == ([n] if k == 1 else [k, n])


@pytest.mark.parametrize("which", ["SA", "LA", "LM"])
@pytest.mark.parametrize("ncv", [0, 9, 14])
def test_thick_restart_with_a_small_basis_vs_eigh(which, ncv):
    """ncv basis vectors (0: the reference's default 2k + 10) far below n: the run restarts
    from the wanted Ritz vectors many times and still returns eigh's eigenpairs."""
    from cytnx_amd.lanczos import _matvec_of, eigsh_native
    n, k = 300, 4
    H = sym(n, 500 + ncv)
    w = np.linalg.eigvalsh(H)
    want = order(w, which)[:k]
    x0 = dev_vec(np.random.default_rng(8).standard_normal(n))
    fn_ctx, keep = _matvec_of(DenseOp(H), x0)
    vals, vecs, matvecs, nconv = eigsh_native(None, x0.storage(), which, k, 0.0, True,
                                              fn_ctx=fn_ctx, Maxiter=5000, ncv=ncv)
    m = ncv or 2 * k + 10
    print(which, "ncv", m, "matvecs", matvecs, "nconv", nconv, flush=True)
    assert nconv == k and matvecs > m  # converged, after at least one restart
    scale = max(1.0, np.abs(w).max())
    assert np.allclose(vals, want, atol=1e-10 * scale, rtol=0)
    V = vecs.cpu().numpy()
    for i in range(k):
        assert abs(np.linalg.norm(V[i]) - 1) < 1e-10
        assert np.linalg.norm(H @ V[i] - vals[i] * V[i]) < 1e-8 * scale
    assert np.abs(V @ V.T - np.eye(k)).max() < 1e-9
    del keep


def test_which_sa_on_a_native_two_site_chain_vs_eigh():
    from cytnx_amd import heisenberg
    from cytnx_amd.heff import Heff2
    psi, L, M1, M2, R = [t.to(0) for t in heisenberg.two_site_problem(8, 11, True)]
    H = Heff2(psi, L, M1, M2, R)
    n = psi._storage.numel()
    cols = []
    for j in range(n):
        e = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        e[j] = 1
        y = torch.empty_like(e)
        H.apply_flat(e, y)
        cols.append(y.cpu().numpy())
    Hm = np.stack(cols, axis=1)
    w = np.linalg.eigvalsh(0.5 * (Hm + Hm.T))
    out = Lanczos(H, psi, which="SA", k=3)
    vals = out[0].storage().cpu().numpy()
    scale = max(1.0, np.abs(w).max())
    assert np.allclose(vals, w[:3], atol=1e-9 * scale, rtol=0)
    assert len(out) == 4  # eigenvalues, then one UniTensor per eigenvector
    for i in range(3):
        x = out[1 + i]._storage.cpu().numpy()
        assert np.linalg.norm(Hm @ x - vals[i] * x) < 1e-7 * scale
    E0, _, _, _, _ = lanczos_gnd_native(H, psi._storage, 1e-13, 400, reorth=True)
    assert abs(vals[0] - E0) < 1e-9 * scale


def test_which_rejects_bad_arguments():
    op = DenseOp(sym(8, 1))
    x0 = dev_vec(np.ones(8))
    with pytest.raises(ValueError):
        Lanczos(op, x0, which="XX")
    with pytest.raises(ValueError):
        Lanczos(op, x0, which="SA", k=9)
    with pytest.raises(ValueError):
        Lanczos(op, x0, which="SA", k=2, ncv=3)


class SwapOp(LinOp):
    def __init__(self, n, coupling=1.0):
        super().__init__("mv", n, 3, 0)
        self.c = coupling

    def matvec(self, v):
        x = v.storage().reshape(-1)
        y = torch.zeros_like(x)
        y[0] = self.c * x[1]
        y[1] = self.c * x[0]
        return Tensor.from_storage(y, [x.numel()])


class ScaleOp(LinOp):
    def __init__(self):
        super().__init__("mv", 1, 3, 0)

    def matvec(self, v):
        x = v.storage().reshape(-1)
        return Tensor.from_storage(3.0 * x, [1])


@pytest.mark.parametrize("tau", [0.1, -0.7, 1.3])
def test_exp_two_dim_swap_closed_form(tau):
    out = Lanczos_Exp(SwapOp(2), dev_vec([1.0, 0.0]), tau, 1e-12)
    assert np.allclose(out.storage().cpu().numpy(), [np.cosh(tau), np.sinh(tau)], atol=1e-12)


@pytest.mark.parametrize("coupling", [1.0, 1e-3, 0.5])
def test_exp_three_dim_with_zero_row_closed_form(coupling):
    tau = 0.4
    out = Lanczos_Exp(SwapOp(3, coupling), dev_vec([1.0, 0.0, 0.0]), tau, 1e-12)
    want = [np.cosh(coupling * tau), np.sinh(coupling * tau), 0.0]
    assert np.allclose(out.storage().cpu().numpy(), want, atol=1e-12)


def test_exp_one_dim_scale():
    out = Lanczos_Exp(ScaleOp(), dev_vec([0.7]), 0.25, 1e-12)
    assert abs(out.storage().item() - 0.7 * np.exp(0.75)) < 1e-12


@pytest.mark.parametrize("n,tau,norm", [(50, 0.1, 1.0), (50, -0.3, 1.1), (200, 0.05, 3.0)])
def test_exp_vs_scipy_expm(n, tau, norm):
    from scipy.linalg import expm
    H = sym(n, n)
    x = np.random.default_rng(n + 1).standard_normal(n)
    x *= norm / np.linalg.norm(x)
    out = Lanczos_Exp(DenseOp(H), dev_vec(x), tau, 1e-13)
    want = expm(tau * H) @ x
    assert np.linalg.norm(out.storage().cpu().numpy() - want) < 1e-10 * max(1.0, norm)


def test_exp_rejects_bad_arguments():
    op = SwapOp(2)
    with pytest.raises(ValueError):
        Lanczos_Exp(op, dev_vec([1.0, 0.0]), 0.1, 0.0)
    with pytest.raises(ValueError):
        Lanczos_Exp(op, dev_vec([1.0, 0.0]), 0.1, 1e-10, 1)
    with pytest.raises(ValueError):
        Lanczos_Exp(op, dev_vec([1.0, 0.0]), complex(float("nan"), 1.0), 1e-10)


@pytest.mark.parametrize("tau", [-0.01, 0.005])
def test_exp_on_a_native_two_site_chain_vs_expm(tau):
    """Lanczos_Exp with the planned chain as the matvec (no Python in the iteration) against
    scipy.linalg.expm of the same operator's dense matrix."""
    from scipy.linalg import expm
    from cytnx_amd import heisenberg
    from cytnx_amd.heff import Heff2
    psi, L, M1, M2, R = [t.to(0) for t in heisenberg.two_site_problem(8, 5, True)]
    H = Heff2(psi, L, M1, M2, R)
    n = psi._storage.numel()
    cols = []
    for j in range(n):
        e = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        e[j] = 1
        y = torch.empty_like(e)
        H.apply_flat(e, y)
        cols.append(y.cpu().numpy())
    Hm = np.stack(cols, axis=1)
    x = psi._storage.cpu().numpy()
    out = Lanczos_Exp(H, psi, tau, 1e-13)
    want = expm(tau * 0.5 * (Hm + Hm.T)) @ x
    got = out._storage.cpu().numpy()
    assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)
