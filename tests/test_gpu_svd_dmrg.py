"""GPU checks of the UniTensor SVDs and the two-site U(1) DMRG.

SVD (block and dense), against what the mathematics fixes:
* U.S.vT reproduces the tensor (1e-12), U and vT^T have orthonormal columns per sector, the
  singular values are those of the dense matrix of the tensor (the block structure only permutes
  it into block-diagonal form), sector keys ascend;
* Svd_truncate keeps the keepdim largest values over all sectors, the Frobenius error of the
  truncated product is the root sum of squares of the dropped values (Eckart-Young per sector),
  and return_err reports the largest / all dropped values.

DMRG, against closed forms and exact diagonalisation:
* the open Heisenberg chain at L=10 and chi=32 (an exact MPS) against the exact ground energy in
  the Sz=0 sector (numpy.linalg.eigvalsh of the 252-state sector matrix), within 1e-9;
* the open XX chain (Jz=0) at L=20, chi=32 against the free-fermion energy
  sum_{k: cos<0} cos(pi k/(L+1)), within 1e-6 (truncated MPS);
* FP32 + FP64 refinement against FP64 on the same run, final energies within 1e-9.
"""
import numpy as np
import pytest
import torch

from cytnx_amd import linalg
from cytnx_amd.dmrg import DMRG
from cytnx_amd.heisenberg import random_block
from cytnx_amd.tensor import Tensor
from cytnx_amd.unitensor import BD_BRA, BD_KET, Bond, Contract, UniTensor

pytestmark = pytest.mark.gpu


def _by_key(T):
    return {tuple(t): T.get_block_(i).detach().cpu().numpy() for i, t in enumerate(T.itoi())}


def _rel(got, want):
    assert sorted(got) == sorted(want)
    num = sum(np.linalg.norm((got[k] - want[k]).ravel()) ** 2 for k in want)
    den = sum(np.linalg.norm(want[k].ravel()) ** 2 for k in want)
    return (num / den) ** 0.5


def _psi_like(chi, seed, complex_=False):
    rng = np.random.default_rng(seed)
    vl = Bond(BD_KET, [[-2], [0], [2]], [chi // 4, chi // 3, chi // 4])
    vr = Bond(BD_KET, [[x] for x in [-4, -2, 0, 2, 4]], [1, chi // 4, chi // 3, chi // 4, 2])
    bp = Bond(BD_KET, [[1], [-1]], [1, 1])
    T = random_block([vl, bp.redirect(), bp.redirect(), vr.redirect()], ["a", "s", "t", "b"],
                     2, rng, np.complex128 if complex_ else np.float64)
    return T.to(0)


def _dense_matrix(T):
    D = T.to_dense()
    r = T.rowrank()
    return D.reshape(int(np.prod(D.shape[:r])), -1)


@pytest.mark.parametrize("rowrank", [1, 2, 3])
@pytest.mark.parametrize("complex_", [False, True])
def test_block_gesvd_reconstructs_and_is_orthonormal(rowrank, complex_):
    T = _psi_like(24, 3 + rowrank, complex_)
    T.set_rowrank_(rowrank)
    S, U, vT = linalg.Gesvd(T)
    assert S.is_diag() and S.labels() == ["_aux_L", "_aux_R"]
    assert U.labels() == T.labels()[:rowrank] + ["_aux_L"] and U.rowrank() == rowrank
    assert vT.labels() == ["_aux_R"] + T.labels()[rowrank:] and vT.rowrank() == 1
    keys = S.bonds()[0].qnums()
    assert keys == sorted(keys)
    R = Contract(Contract(U, S), vT)
    assert R.labels() == T.labels()
    assert _rel(_by_key(R), _by_key(T)) <= 1e-12
    # singular values: those of the dense matrix
    sv = np.sort(S.storage().cpu().numpy())[::-1]
    ref = np.linalg.svd(_dense_matrix(T), compute_uv=False)
    np.testing.assert_allclose(sv, ref[:sv.size], rtol=1e-12, atol=1e-12 * ref[0])
    assert np.all(ref[sv.size:] <= 1e-12 * ref[0])
    # orthonormal columns: U^H U = I and vT vT^H = I on the dense forms
    Ud = U.to_dense().reshape(-1, U.shape()[-1])
    Vd = vT.to_dense().reshape(vT.shape()[0], -1)
    k = Ud.shape[1]
    np.testing.assert_allclose(Ud.conj().T @ Ud, np.eye(k), atol=1e-12)
    np.testing.assert_allclose(Vd @ Vd.conj().T, np.eye(k), atol=1e-12)


@pytest.mark.parametrize("keep", [1, 7, 30])
def test_block_svd_truncate_eckart_young(keep):
    T = _psi_like(32, 11)
    S0 = np.sort(linalg.Gesvd(T, False, False)[0].storage().cpu().numpy())[::-1]
    S, U, vT, err = linalg.Svd_truncate(T, keep, return_err=2)
    kept = np.sort(S.storage().cpu().numpy())[::-1]
    np.testing.assert_allclose(kept, S0[:keep], rtol=1e-12)
    dropped = S0[keep:]
    np.testing.assert_allclose(err.storage().cpu().numpy(), dropped, rtol=1e-12)
    R = Contract(Contract(U, S), vT)
    diff = np.linalg.norm(R.to_dense() - T.to_dense())
    assert abs(diff - np.sqrt(np.sum(dropped ** 2))) <= 1e-10 * np.linalg.norm(S0)
    e1 = linalg.Svd_truncate(T, keep, return_err=1)[3]
    assert e1.item() == pytest.approx(dropped[0], rel=1e-12)


def test_block_svd_truncate_err_threshold():
    T = _psi_like(32, 12)
    S0 = np.sort(linalg.Gesvd(T, False, False)[0].storage().cpu().numpy())[::-1]
    thr = 0.5 * (S0[5] + S0[6])
    S = linalg.Svd_truncate(T, 100, err=thr)[0]
    assert S.storage().numel() == 6


@pytest.mark.parametrize("absorb", ["right", "left"])
@pytest.mark.parametrize("complex_", [False, True])
def test_truncated_split_matches_svd_truncate(absorb, complex_):
    """The density-matrix split: the same kept values and the same rank-k product as
    Svd_truncate, the kept factor orthonormal to working precision."""
    from cytnx_amd.svd import truncated_split
    T = _psi_like(40, 21, complex_)
    keep = 17
    S, U, vT = linalg.Svd_truncate(T, keep)
    S2, X, Y = truncated_split(T, keep, absorb)
    np.testing.assert_allclose(np.sort(S2.storage().cpu().numpy()),
                               np.sort(S.storage().cpu().numpy()), rtol=1e-10)
    want = Contract(Contract(U, S), vT).to_dense()
    got = Contract(X, Y)
    assert got.labels() == T.labels()
    assert np.linalg.norm(got.to_dense() - want) <= 1e-10 * np.linalg.norm(want)
    if absorb == "right":
        assert X.labels()[-1] == "_aux_L" and Y.labels()[0] == "_aux_L"
        Q = X.to_dense().reshape(-1, X.shape()[-1])
        np.testing.assert_allclose(Q.conj().T @ Q, np.eye(Q.shape[1]), atol=1e-13)
    else:
        assert X.labels()[-1] == "_aux_R" and Y.labels()[0] == "_aux_R"
        Q = Y.to_dense().reshape(Y.shape()[0], -1)
        np.testing.assert_allclose(Q @ Q.conj().T, np.eye(Q.shape[0]), atol=1e-13)


@pytest.mark.parametrize("rowrank", [1, 2])
@pytest.mark.parametrize("complex_", [False, True])
def test_orthogonal_split_reconstructs(rowrank, complex_):
    from cytnx_amd.svd import orthogonal_split
    T = _psi_like(40, 31 + rowrank, complex_)
    T.set_rowrank_(rowrank)
    X, Q = orthogonal_split(T)
    S = linalg.Gesvd(T, False, False)[0]
    assert Q.bonds()[0].qnums() == S.bonds()[0].qnums()
    assert Q.bonds()[0].getDegeneracies() == S.bonds()[0].getDegeneracies()
    R = Contract(X, Q)
    assert R.labels() == T.labels()
    assert _rel(_by_key(R), _by_key(T)) <= 1e-12
    Qd = Q.to_dense().reshape(Q.shape()[0], -1)
    np.testing.assert_allclose(Qd @ Qd.conj().T, np.eye(Qd.shape[0]), atol=1e-13)


def test_dense_svd_truncate():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((6, 5, 7))
    T = UniTensor([Bond(6), Bond(5), Bond(7)], ["a", "b", "c"], 2, 3, -1)
    T._block = Tensor.from_numpy(a, device=0)
    S, U, vT = linalg.Svd_truncate(T, 4)
    ref = np.linalg.svd(a.reshape(30, 7), compute_uv=False)
    np.testing.assert_allclose(S.get_block_().cpu().numpy(), ref[:4], rtol=1e-12)
    assert U.shape() == [6, 5, 4] and vT.shape() == [4, 7]
    R = Contract(Contract(U, S), vT)
    assert abs(np.linalg.norm(R.to_dense() - a) - np.sqrt(np.sum(ref[4:] ** 2))) <= 1e-10


def _heisenberg_ed(L, J=1.0, Jz=1.0):
    """Ground energy of the open XXZ chain in the Sz=0 sector (dense eigvalsh)."""
    states = [s for s in range(1 << L) if bin(s).count("1") == L // 2]
    idx = {s: i for i, s in enumerate(states)}
    H = np.zeros((len(states), len(states)))
    for s in states:
        i = idx[s]
        for b in range(L - 1):
            u, v = (s >> b) & 1, (s >> (b + 1)) & 1
            H[i, i] += Jz * (0.25 if u == v else -0.25)
            if u != v:
                t = s ^ (1 << b) ^ (1 << (b + 1))
                H[idx[t], i] += 0.5 * J
    return np.linalg.eigvalsh(H)[0]


def test_dmrg_heisenberg_exact_at_L10():
    dm = DMRG(10, 32, seed=3, maxit=30)
    E = dm.run(3)
    ref = _heisenberg_ed(10)
    print("DMRG L=10 chi=32: E", E[-1], "ED", ref, "bond dims", dm.bond_dims(), flush=True)
    assert abs(E[-1] - ref) <= 1e-9


def test_dmrg_xx_free_fermions_at_L20():
    L = 20
    dm = DMRG(L, 32, Jz=0.0, seed=5, maxit=30)
    E = dm.run(3)
    c = np.cos(np.pi * np.arange(1, L + 1) / (L + 1))
    ref = c[c < 0].sum()
    print("DMRG XX L=20 chi=32: E", E[-1], "free fermions", ref, flush=True)
    assert abs(E[-1] - ref) <= 1e-6


def test_dmrg_density_matrix_split_matches_gesvd_split():
    a = DMRG(12, 24, seed=9, maxit=30, svd="gesvd")
    b = DMRG(12, 24, seed=9, maxit=30, svd="gram")
    Ea = a.run(2)
    Eb = b.run(2)
    print("DMRG L=12 chi=24 gesvd", Ea[-1], "gram", Eb[-1], flush=True)
    assert abs(Ea[-1] - Eb[-1]) <= 1e-10
    assert a.bond_dims() == b.bond_dims()


def test_dmrg_mixed_precision_matches_fp64():
    a = DMRG(12, 32, seed=7, maxit=30, prec="fp64")
    b = DMRG(12, 32, seed=7, maxit=30, prec="mixed")
    Ea = a.run(2)
    Eb = b.run(2)
    print("DMRG L=12 chi=32 fp64", Ea[-1], "mixed", Eb[-1], flush=True)
    assert abs(Ea[-1] - Eb[-1]) <= 1e-9
    assert abs(Ea[-1] - _heisenberg_ed(12)) <= 1e-6


def test_random_right_canonical_mps():
    """The DMRG starting state: every A_k (k >= 1) a right isometry (A A^T = I over (s, a_{k+1}))
    and A_0 of unit norm, with the bond dimensions the right-to-left LQ sweep of random_mps
    produces."""
    from cytnx_amd.dmrg import random_mps, random_right_canonical_mps, right_canonicalize
    L, chi = 12, 16
    As = random_right_canonical_mps(L, chi, seed=4, device=0)
    Bs = right_canonicalize(random_mps(L, chi, 4, 0), "gram")
    assert [A.shape() for A in As] == [B.shape() for B in Bs]
    assert abs(float(As[0].Norm()) - 1.0) <= 1e-13
    for k in range(1, L):
        D = As[k].to_dense()
        M = D.reshape(D.shape[0], -1)
        err = np.abs(M @ M.T - np.eye(M.shape[0])).max()
        print("site", k, "shape", D.shape, "max |A A^T - I|", err, flush=True)
        assert err <= 1e-13
