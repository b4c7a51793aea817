"""Parity at the sizes of the measured workloads.

* Two-site U(1) Heff at chi=1024 (31 Gaussian sectors, seed 1: the benchmark's operator) and at
  chi=2048 (35 sectors, the DMRG centre-bond size), each precision with the tilings the
  autotuner picks, against the oracle's chain of restated BlockUniTensor::contract calls
  (src/BlockUniTensor.cpp:1114-1214) in the reference's order: FP64 within 1e-12 normwise,
  FP32 within 5e-6; FP32 + refinement within 1e-10 of the FP64 Lanczos energy.
* Two-site DMRG of the open Heisenberg chain, L=100, chi=2048, one sweep: the FP64 run against
  the Bethe-ansatz ground energy (tests/bethe.py), and FP32 + refinement against FP64.
* MPS energy gradient at L=32, chi=512 (complex128): central differences along random
  directions, and complex64 against complex128.
"""
import numpy as np
import pytest
import torch

import pyoracle as po
from bethe import open_xxx_ground_energy
from cytnx_amd import heisenberg
from cytnx_amd.heff import Heff2
from cytnx_amd.lanczos import eigsh_native, ground_state_mixed, lanczos_gnd_native
from cytnx_amd.unitensor import PREC_FP32, PREC_FP32X3, PREC_FP64

pytestmark = pytest.mark.gpu


def relerr(got, ref):
    d = np.linalg.norm((np.asarray(got) - np.asarray(ref)).ravel())
    n = np.linalg.norm(np.asarray(ref).ravel())
    return d / n if n else d


def _oblock(ut):
    blocks = [ut.get_block_(i).detach().cpu().numpy() for i in range(ut.Nblocks())]
    bonds = [po.OBond(b.type(), b.qnums(), b.getDegeneracies()) for b in ut.bonds()]
    return po.OBlockTensor(bonds, ut.labels(), ut.rowrank(), itoi=ut.itoi(), blocks=blocks)


def _oracle_chain(host):
    psi, L, M1, M2, R = [_oblock(t) for t in host]
    X = po.block_contract(L, psi)
    X = po.block_contract(X, M1)
    X = po.block_contract(X, M2)
    return po.block_contract(X, R)


@pytest.fixture(scope="module", params=[1024, 2048])
def sized_problem(request):
    chi = request.param
    host = heisenberg.two_site_problem(chi, 1, hermitian=True)
    ref = _oracle_chain(host).flat()
    psi, L, M1, M2, R = [t.to(0) for t in host]
    H = Heff2(psi, L, M1, M2, R)
    assert H.mode == "mix"
    return chi, psi, H, ref


@pytest.mark.parametrize("prec,tol", [(PREC_FP64, 1e-12), (PREC_FP32, 5e-6),
                                      (PREC_FP32X3, 5e-6)])
def test_heff2_tuned_matvec_matches_oracle_at_size(sized_problem, prec, tol):
    chi, psi, H, ref = sized_problem
    H.set_precision(prec)
    choice = H.tune(reps=3)
    y = torch.empty_like(psi._storage)
    H.apply_flat(psi._storage, y)
    err = relerr(y.cpu().numpy(), ref)
    print("chi", chi, "prec", prec, "tiles", choice, "normwise error vs oracle %.3e" % err,
          flush=True)
    H.set_precision(PREC_FP64)
    assert err <= tol


@pytest.mark.parametrize("sized_problem", [1024], indirect=True)
def test_fp32_refinement_energy_at_chi1024(sized_problem):
    chi, psi, H, ref = sized_problem
    H.set_precision(PREC_FP64)
    x0 = psi._storage.clone()
    E64, _, n64, _, _ = lanczos_gnd_native(H, x0, 1e-12, 600, batch=16)
    ev, _, m, nconv = eigsh_native(H, x0, "SA", 1, 1e-12, False)
    Emx, _, n1, n2 = ground_state_mixed(H, x0, 1e-12, 600, batch=16)
    print("chi=1024: E64 %.15f (%d it)  eigsh %.15f (%d vectors)  mixed %.15f (%d + %d it)"
          % (E64, n64, ev[0], m, Emx, n1, n2), flush=True)
    assert nconv == 1
    assert abs(E64 - ev[0]) <= 1e-9 * abs(ev[0])
    assert abs(Emx - E64) <= 1e-10


def test_dmrg_L100_chi2048_against_bethe_ansatz():
    from cytnx_amd.dmrg import DMRG
    Eb = open_xxx_ground_energy(100)
    a = DMRG(100, 2048, seed=1, maxit=30, prec="fp64")
    Ea = a.run(1)[-1]
    b = DMRG(100, 2048, seed=1, maxit=30, prec="mixed")
    Emx = b.run(1)[-1]
    print("L=100 chi=2048, one sweep: fp64 %.14f mixed %.14f Bethe %.14f max bond %d"
          % (Ea, Emx, Eb, max(a.bond_dims())), flush=True)
    assert max(a.bond_dims()) == 2048
    assert abs(Ea - Eb) <= 1e-9
    assert abs(Emx - Ea) <= 1e-10


def test_mps_gradient_at_L32_chi512():
    from cytnx_amd.mps import mps_energy, random_mps
    As = random_mps(32, 512, seed=1, dtype=torch.complex128)
    assert max(max(A.shape()) for A in As) == 512
    for A in As:
        A._storage.requires_grad_(True)
    E, _, _ = mps_energy(As, checkpoint=True)
    E.backward()
    grads = [A._storage.grad.clone() for A in As]
    rng = np.random.default_rng(32)
    h = 1e-6
    for trial in range(2):
        D = [torch.from_numpy(rng.standard_normal(A._storage.numel()) +
                              1j * rng.standard_normal(A._storage.numel())).cuda() for A in As]
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
        print("L=32 chi=512 trial", trial, "fd %.12e analytic %.12e" % (fd, an), flush=True)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an))
    # scale invariance of E = <psi|H|psi>/<psi|psi>: Re <A, dE/dA*> = 0 for every site
    for A, g in zip(As, grads):
        assert abs(torch.sum(torch.conj(g) * A._storage.detach()).real.item()) <= \
            1e-8 * torch.linalg.vector_norm(g).item() * torch.linalg.vector_norm(A._storage).item()
    # complex64 against complex128
    As32 = []
    for A in As:
        B = A.clone_meta()
        B._storage = A._storage.detach().to(torch.complex64).requires_grad_(True)
        As32.append(B)
    E32, _, _ = mps_energy(As32)
    E32.backward()
    num = sum(torch.linalg.vector_norm(B._storage.grad.to(torch.complex128) - g).item() ** 2
              for B, g in zip(As32, grads))
    den = sum(torch.linalg.vector_norm(g).item() ** 2 for g in grads)
    print("c64 vs c128: dE %.3e grad rel err %.3e" % (E32.item() - E.item(), (num / den) ** 0.5),
          flush=True)
    assert abs(E32.item() - E.item()) <= 1e-4 * abs(E.item())
    assert (num / den) ** 0.5 <= 1e-3
