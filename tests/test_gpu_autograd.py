"""GPU checks of the differentiable block contraction and the MPS energy gradient.

* contract_ad: gradients of a random real loss Re(sum(conj(R) * C)) against torch's CPU FP64
  autograd on the dense forms of the same tensors (entries inside the blocks), real and
  complex, with common legs in scattered positions so the inverse permutes matter.
* mps_energy: the energy against a dense CPU contraction of the same MPS and MPO at small L,
  its gradient against dense CPU autograd, central finite differences (h = 1e-6) along random
  directions at a larger size, recomputed (checkpointed) environments giving the same gradient,
  and the complex64 gradient against complex128.
"""
import numpy as np
import pytest
import torch

from cytnx_amd import heisenberg
from cytnx_amd.autograd import contract_ad
from cytnx_amd.mps import mps_energy, random_mps
from cytnx_amd.unitensor import BD_BRA, BD_KET, Bond, UniTensor

pytestmark = pytest.mark.gpu


def _offs(T):
    return [np.concatenate([[0], np.cumsum(b.getDegeneracies())[:-1]]).astype(int)
            for b in T.bonds()]


def _dense_positions(T):
    """(block slices) for mapping dense arrays to block storage order."""
    offs = _offs(T)
    out = []
    for bi, t in enumerate(T._itoi):
        shp = [int(s) for s in T._shapes[bi]]
        out.append(tuple(slice(offs[a][int(q)], offs[a][int(q)] + shp[a])
                         for a, q in enumerate(t)))
    return out


def _to_flat(T, D):
    return np.concatenate([np.asarray(D[sl]).ravel() for sl in _dense_positions(T)])


def _rand_block(bonds, labels, rng, complex_):
    T = UniTensor(bonds, labels, 1, 1 if complex_ else 3, -1)
    n = T.storage().numel()
    x = rng.standard_normal(n) + (1j * rng.standard_normal(n) if complex_ else 0)
    T._storage = torch.from_numpy(np.asarray(x))
    return T.to(0)


@pytest.mark.parametrize("complex_", [False, True])
def test_contract_ad_matches_dense_autograd(complex_):
    rng = np.random.default_rng(7 + complex_)
    b1 = Bond(BD_KET, [[-2], [0], [2]], [3, 4, 2])
    b2 = Bond(BD_KET, [[-1], [1]], [2, 3])
    b3 = Bond(BD_KET, [[-1], [0], [1]], [2, 1, 3])
    b4 = Bond(BD_KET, [[0], [2]], [3, 2])
    bl = Bond(BD_KET, [[-1], [1], [3]], [2, 3, 1])
    # A(i, k, j, m) and B(m, l, k): common legs k and m sit in the middle / at the ends
    A = _rand_block([b1, b2, b3.redirect(), b4], ["i", "k", "j", "m"], rng, complex_)
    B = _rand_block([b4.redirect(), bl, b2.redirect()], ["m", "l", "k"], rng, complex_)
    assert A.Nblocks() > 1 and B.Nblocks() > 1
    A._storage.requires_grad_(True)
    B._storage.requires_grad_(True)
    C = contract_ad(A, B)
    assert C.labels() == ["i", "j", "l"]
    Rw = torch.from_numpy(rng.standard_normal(C._storage.numel()) +
                          (1j * rng.standard_normal(C._storage.numel()) if complex_ else 0)).cuda()
    loss = torch.sum(torch.conj(Rw) * C._storage).real if complex_ else torch.sum(Rw * C._storage)
    loss.backward()
    gA = A._storage.grad.cpu().numpy()
    gB = B._storage.grad.cpu().numpy()
    # dense CPU reference with the same entries
    dt = torch.complex128 if complex_ else torch.float64
    Ad = torch.tensor(A.to_dense(), dtype=dt, requires_grad=True)
    Bd = torch.tensor(B.to_dense(), dtype=dt, requires_grad=True)
    Cd = torch.einsum("ikjm,mlk->ijl", Ad, Bd)
    Rd = np.zeros(Cd.shape, dtype=np.complex128 if complex_ else np.float64)
    Cm = C.clone_meta()
    for sl, (bi, _) in zip(_dense_positions(Cm), enumerate(Cm._itoi)):
        o = int(Cm._offsets[bi])
        n = int(np.prod(Cm._shapes[bi].astype(np.int64)))
        Rd[sl] = Rw.cpu().numpy()[o:o + n].reshape(Rd[sl].shape)
    Rt = torch.from_numpy(Rd)
    ld = torch.sum(torch.conj(Rt) * Cd).real if complex_ else torch.sum(Rt * Cd)
    ld.backward()
    np.testing.assert_allclose(C._storage.detach().cpu().numpy(), _to_flat(C, Cd.detach().numpy()),
                               rtol=1e-12, atol=1e-12)
    wantA = _to_flat(A, Ad.grad.numpy())
    wantB = _to_flat(B, Bd.grad.numpy())
    assert np.linalg.norm(gA - wantA) <= 1e-12 * np.linalg.norm(wantA)
    assert np.linalg.norm(gB - wantB) <= 1e-12 * np.linalg.norm(wantB)


def _dense_energy(As, dt=torch.complex128):
    """<psi|H|psi> / <psi|psi> on the CPU with dense site tensors (requires_grad)."""
    W = torch.tensor(heisenberg.xxz_mpo_block().to_dense(), dtype=dt)  # grouped state order
    Ds = [torch.tensor(A.to_dense(), dtype=dt, requires_grad=True) for A in As]
    E = torch.zeros((1, 5, 1), dtype=dt)
    E[0, 2, 0] = 1.0  # left boundary: state 4 = grouped index 2
    F = torch.ones((1, 1), dtype=dt)
    for A in Ds:
        E = torch.einsum("awb,asx,wvst,bty->xvy", E, A, W, torch.conj(A))
        F = torch.einsum("ab,asx,bsy->xy", F, A, torch.conj(A))
    num = E[0, 0, 0].real  # right boundary: state 0 = grouped index 0
    den = F.sum().real
    return num / den, Ds


def test_mps_energy_and_gradient_match_dense():
    As = random_mps(6, 8, seed=3, dtype=torch.complex128)
    for A in As:
        A._storage.requires_grad_(True)
    E, num, den = mps_energy(As)
    E.backward()
    Ed, Ds = _dense_energy([_cpu_copy(A) for A in As])
    Ed.backward()
    assert abs(E.item() - Ed.item()) <= 1e-12 * abs(Ed.item())
    for A, D in zip(As, Ds):
        want = _to_flat(A, D.grad.numpy())
        got = A._storage.grad.cpu().numpy()
        assert np.linalg.norm(got - want) <= 1e-10 * np.linalg.norm(want)


def _cpu_copy(A):
    B = A.clone_meta()
    B._storage = A._storage.detach().cpu()
    return B


def test_mps_gradient_finite_differences_and_checkpoint():
    L, chi = 12, 32
    As = random_mps(L, chi, seed=5, dtype=torch.complex128)
    for A in As:
        A._storage.requires_grad_(True)
    E, _, _ = mps_energy(As)
    E.backward()
    grads = [A._storage.grad.clone() for A in As]
    # recomputed environments: the same gradient
    for A in As:
        A._storage.grad = None
    E2, _, _ = mps_energy(As, checkpoint=True)
    E2.backward()
    assert E2.item() == E.item()
    for A, g in zip(As, grads):
        assert torch.equal(A._storage.grad, g)
    # central differences along random complex directions: dE = Re(sum(conj(g) * dA))
    rng = np.random.default_rng(11)
    h = 1e-6
    for trial in range(3):
        D = [torch.from_numpy(rng.standard_normal(A._storage.numel()) +
                              1j * rng.standard_normal(A._storage.numel())).cuda()
             for A in As]
        with torch.no_grad():
            def E_at(sign):
                Bs = []
                for A, d in zip(As, D):
                    B = A.clone_meta()
                    B._storage = A._storage.detach() + sign * h * d
                    Bs.append(B)
                return mps_energy(Bs)[0].item()
            fd = (E_at(1) - E_at(-1)) / (2 * h)
        an = sum(torch.sum(torch.conj(g) * d).real.item() for g, d in zip(grads, D))
        print("trial", trial, "fd", fd, "analytic", an, flush=True)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))


def test_mps_gradient_complex64_close_to_complex128():
    As64 = random_mps(10, 24, seed=9, dtype=torch.complex128)
    As32 = []
    for A in As64:
        B = A.clone_meta()
        B._storage = A._storage.to(torch.complex64)
        As32.append(B)
    for A in As64 + As32:
        A._storage.requires_grad_(True)
    E64, _, _ = mps_energy(As64)
    E64.backward()
    E32, _, _ = mps_energy(As32)
    E32.backward()
    assert abs(E32.item() - E64.item()) <= 1e-4 * abs(E64.item())
    num = sum(torch.linalg.vector_norm(a.grad - b.grad.to(torch.complex128)).item() ** 2
              for a, b in zip([A._storage for A in As64], [B._storage for B in As32]))
    den = sum(torch.linalg.vector_norm(a.grad).item() ** 2 for a in [A._storage for A in As64])
    assert (num / den) ** 0.5 <= 1e-3
