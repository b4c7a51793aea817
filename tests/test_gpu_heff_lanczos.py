"""GPU parity of the effective-Hamiltonian chains and the Lanczos driver.

* Heff2 (two-site U(1)): every block against the oracle's chain of restated
  BlockUniTensor::contract calls in the same order, and against the dense closed form at small
  chi; FP64 within 1e-12 normwise, FP32 / FP32X3 error levels bounded and recorded.
* Heff1Dense (one-site, C1 layout): against the oracle's tensordot chain at chi=128.
* Lanczos: the reference's known 27x27 eigenvalue (-1628.9964650426593 within 1e-5), the
  lan_block_{A,B,C} U(1) operator (eigenvalue within 1e-12 of eigh, residual < 1e-6), Heff2 at
  small chi against numpy.linalg.eigh of the explicit matrix, FP32+refinement within 1e-10 of
  FP64, and graph replay identical to plain launches.
* Set-up reuse: Lanczos contexts reshaped across vector lengths give bit-identical runs, and
  chains built from pooled device blocks still match the oracle."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import REF_FIXTURES
from cytnx_amd import fileio, heisenberg, linalg
from cytnx_amd.heff import Heff1Dense, Heff2
from cytnx_amd.lanczos import Lanczos, LinOp, ground_state_mixed, lanczos_gnd_native
from cytnx_amd.tensor import Tensor
from cytnx_amd.unitensor import BD_BRA, BD_KET, Bond, UniTensor

pytestmark = pytest.mark.gpu


def relerr(got, ref):
    d = np.linalg.norm((np.asarray(got) - np.asarray(ref)).ravel())
    n = np.linalg.norm(np.asarray(ref).ravel())
    return d / n if n else d


def _oblock(ut):
    blocks = [ut.get_block_(i).detach().cpu().numpy() for i in range(ut.Nblocks())]
    bonds = [po.OBond(b.type(), b.qnums(), b.getDegeneracies()) for b in ut.bonds()]
    return po.OBlockTensor(bonds, ut.labels(), ut.rowrank(), itoi=ut.itoi(), blocks=blocks)


def oracle_heff2(psi, L, M1, M2, R):
    X = po.block_contract(_oblock(L), _oblock(psi))
    Y = po.block_contract(X, _oblock(M1))
    Z = po.block_contract(Y, _oblock(M2))
    return po.block_contract(Z, _oblock(R))


def problem(chi, seed=1, hermitian=True):
    psi, L, M1, M2, R = heisenberg.two_site_problem(chi, seed, hermitian)
    return [t.to(0) for t in (psi, L, M1, M2, R)]


@pytest.mark.parametrize("mode", ["ref", "w12", "mix"])
@pytest.mark.parametrize("chi", [16, 64, 160])
def test_heff2_fp64_matches_oracle_chain(chi, mode):
    psi, L, M1, M2, R = problem(chi, seed=chi)
    H = Heff2(psi, L, M1, M2, R, fuse_mpo=mode != "ref", mix=mode == "mix")
    print("heff2 chi", chi, "mode", H.mode, "pairs", H.pairs, "copy bytes", H.copies, flush=True)
    assert H.mode == mode
    assert all(c == 0 for c in H.copies)  # every operand read in place
    out = H.matvec(psi)
    torch.cuda.synchronize()
    ref = oracle_heff2(psi, L, M1, M2, R)
    assert ref.itoi == psi.itoi()
    got = out._storage.cpu().numpy()
    assert relerr(got, ref.flat()) <= 1e-12
    if mode == "mix":
        # both fixed operands read k4-packed, Y written packed by the mix and read packed by R
        assert [H.chain.layout(k) for k in range(2)] == [1 | 4, 1 | 2]


def test_heff2_dense_closed_form_and_symmetry():
    psi, L, M1, M2, R = problem(8, seed=3)
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
    # closed form: dense contraction of the dense tensors, restricted to psi's blocks
    Ld, M1d, M2d, Rd = (t.to_dense() for t in (L, M1, M2, R))
    full = np.einsum("wab,wvsx,vuty,ucd->bxydastc", Ld, M1d, M2d, Rd, optimize=True)
    D = psi.shape()
    full = full.reshape(np.prod(D), np.prod(D))
    # map block-flat positions to dense positions
    pos = []
    offs = [np.concatenate([[0], np.cumsum(b.getDegeneracies())[:-1]]).astype(int)
            for b in psi.bonds()]
    for bi, t in enumerate(psi.itoi()):
        shp = [int(s) for s in psi._shapes[bi]]
        idx = np.indices(shp).reshape(len(shp), -1).T
        for mi in idx:
            g = [offs[a][t[a]] + mi[a] for a in range(4)]
            pos.append(np.ravel_multi_index(g, D))
    sub = full[np.ix_(pos, pos)]
    assert relerr(Hm, sub) <= 1e-12
    assert np.abs(Hm - Hm.T).max() <= 1e-12 * np.abs(Hm).max()  # hermitised environments


@pytest.mark.parametrize("mix", [True, False])
def test_heff2_tuned_tilings_match_oracle_at_chi512(mix):
    """After the autotuner picks tilings (MFMA shapes, split-K, row streaming) for FP64, the
    matvec still equals the oracle's chain at 1e-12."""
    psi, L, M1, M2, R = problem(512, seed=21)
    H = Heff2(psi, L, M1, M2, R, mix=mix)
    choice = H.tune(reps=3)
    print("tuned", H.mode, choice, flush=True)
    out = H.matvec(psi)
    ref = oracle_heff2(psi, L, M1, M2, R)
    assert relerr(out._storage.cpu().numpy(), ref.flat()) <= 1e-12


@pytest.mark.parametrize("split", [0, 1])
@pytest.mark.parametrize("kind", [0, 2, 3, 1])
def test_heff2_mix_every_tile_kind_matches_oracle(kind, split):
    """Each tile shape of the mix launch (16x16, 32x16, 16x32, 32x32) with eight-wave (one
    slice per wave) and four-wave (two slices per wave) workgroups, at chi=256, where sector
    widths (up to 21) leave partial tiles in both directions."""
    psi, L, M1, M2, R = problem(256, seed=23)
    H = Heff2(psi, L, M1, M2, R)
    assert H.mode == "mix"
    H.set_tiles([{"kind": kind, "split": split}, {"kind": -1, "split": -1}])
    out = H.matvec(psi)
    ref = oracle_heff2(psi, L, M1, M2, R)
    assert relerr(out._storage.cpu().numpy(), ref.flat()) <= 1e-12


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 5e-6)])
def test_heff2_R_step_every_tiling_matches_oracle(prec, tol):
    """The grouped GEMM of the .R launch with every MFMA tile shape and split-K mode (none,
    four-wave, eight-wave workgroups) and the 64x32 shape, at chi=256."""
    psi, L, M1, M2, R = problem(256, seed=29)
    H = Heff2(psi, L, M1, M2, R)
    H.set_precision(prec)
    ref = oracle_heff2(psi, L, M1, M2, R).flat()
    cases = [(k, s) for k in (0, 1, 2, 3) for s in (0, 1, 2)] + [(4, 0)]
    for kind, split in cases:
        H.set_tiles([{"kind": -1, "split": -1}, {"kind": kind, "split": split}])
        out = H.matvec(psi)
        err = relerr(out._storage.cpu().numpy(), ref)
        print("R step kind", kind, "split", split, "prec", prec, "err", err, flush=True)
        assert err <= tol


@pytest.mark.parametrize("prec,tol", [(1, 5e-6), (2, 5e-6)])
def test_heff2_reduced_precision_error_levels(prec, tol):
    # chi=384: sectors up to 31 wide, so the MFMA tiles (not only row streaming) run
    psi, L, M1, M2, R = problem(384, seed=5)
    H = Heff2(psi, L, M1, M2, R)
    y64 = H.matvec(psi)._storage.cpu().numpy()
    H.set_precision(prec)
    y = H.matvec(psi)._storage.cpu().numpy()
    err = relerr(y, y64)
    print("precision", prec, "normwise relative error", err, flush=True)
    assert err <= tol


def test_heff1_dense_matches_oracle_c1_shape():
    rng = np.random.default_rng(42)
    chi, D, d = 128, 5, 2
    L = rng.standard_normal((chi, D, chi))
    A = rng.standard_normal((chi, d, chi))
    R = rng.standard_normal((chi, D, chi))
    W = heisenberg.xxz_mpo_dense()
    H = Heff1Dense(torch.from_numpy(L).cuda(), torch.from_numpy(W).cuda(),
                   torch.from_numpy(R).cuda(), A.shape)
    assert H.flops == 4 * D * d * chi ** 3 + 2 * D * D * d * d * chi * chi
    y = torch.empty(A.size, dtype=torch.float64, device="cuda:0")
    H.apply_flat(torch.from_numpy(A.ravel()).cuda(), y)
    X = po.tensordot(L, A, [0], [0])              # (w, l', s, r)
    Y = po.tensordot(X, W, [2, 0], [2, 0])        # (l', r, w', s')
    O = po.tensordot(Y, R, [1, 2], [0, 1])        # (l', s', r')
    assert relerr(y.cpu().numpy(), O.ravel()) <= 1e-12


class DenseOp(LinOp):
    def __init__(self, M):
        super().__init__("mv", M.shape[0], 3, 0)
        self.M = Tensor.from_numpy(M, 0)

    def matvec(self, v):
        out = v.clone_meta()
        x = Tensor.from_storage(v.storage().reshape(-1), [self.nx(), 1])
        out._block = linalg.Matmul(self.M, x)
        return out


def _dense_ut(vec):
    ut = UniTensor([Bond(len(vec)), Bond(1)], ["b", "c"], 1, 3, 0)
    ut._block = Tensor.from_storage(torch.from_numpy(np.asarray(vec, dtype=np.float64)).cuda(),
                                    [len(vec), 1])
    return ut


def test_lanczos_reference_27x27_known_eigenvalue():
    A = np.arange(27 * 27, dtype=np.float64).reshape(27, 27)
    A = A + A.T
    E, v = Lanczos(DenseOp(A), _dense_ut(np.arange(27.0)), "Gnd", 1e-14, 10000)
    assert abs(E.item() - (-1628.9964650426593)) < 1e-5


def test_lanczos_block_operator_fixture():
    blocks = [fileio.read_tensor(os.path.join(REF_FIXTURES, "lan_block_%s.cytn" % n)).logical()
              for n in "ABC"]
    H = np.zeros((27, 27))
    for i, b in enumerate(blocks):
        H[9 * i:9 * i + 9, 9 * i:9 * i + 9] = b
    E0 = min(np.linalg.eigvalsh(b)[0] for b in blocks)
    rng = np.random.default_rng(9)
    E, v = Lanczos(DenseOp(H), _dense_ut(rng.standard_normal(27)), "Gnd", 1e-15, 10000)
    vv = v.storage().reshape(-1).cpu().numpy()
    assert abs(E.item() - E0) < 1e-12
    assert np.linalg.norm(H @ vv - E.item() * vv) < 1e-6
    # the oracle's restated recurrence reaches the same energy
    Eo, _, _ = po.lanczos_gnd(lambda x: H @ x, rng.standard_normal(27), 1e-15, 10000)
    assert abs(Eo - E.item()) < 1e-12


def test_heff2_lanczos_vs_eigh():
    psi, L, M1, M2, R = problem(8, seed=11)
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
    E, v, k, a, b = lanczos_gnd_native(H, psi._storage, 1e-13, 2000, batch=4, reorth=True)
    print("lanczos iterations", k, flush=True)
    assert abs(E - w[0]) <= 1e-10 * max(1, abs(w[0]))
    vv = v.cpu().numpy()
    assert np.linalg.norm(Hm @ vv - E * vv) <= 1e-5 * abs(E)


@pytest.mark.parametrize("prec", [0, 1])
def test_lanczos_batching_and_graph_do_not_change_results(prec):
    psi, L, M1, M2, R = problem(64, seed=13)
    H = Heff2(psi, L, M1, M2, R)
    H.set_precision(prec)
    r1 = lanczos_gnd_native(H, psi._storage, 1e-12, 60, batch=1)
    r2 = lanczos_gnd_native(H, psi._storage, 1e-12, 60, batch=7)
    r3 = lanczos_gnd_native(H, psi._storage, 1e-12, 60, batch=5, graph=True)
    assert r1[2] == r2[2] == r3[2]
    assert r1[0] == r2[0] == r3[0]
    assert np.array_equal(r1[3], r2[3]) and np.array_equal(r1[3], r3[3])
    assert torch.equal(r1[1], r3[1])


@pytest.mark.parametrize("prec", [0, 1])
def test_fused_iteration_matches_two_launch_iteration(prec):
    """A chain matvec runs the fused iteration (the .R launch's epilogue leaves the Hu . u and
    u . u partials, one residual launch follows); the same operator behind a plain callback runs
    the two vector launches.  Same iterations, alphas and betas to rounding, same energy."""
    import ctypes
    from cytnx_amd import _lib
    from cytnx_amd.lanczos import _MATVEC_T
    psi, L, M1, M2, R = problem(96, seed=29)
    H = Heff2(psi, L, M1, M2, R)
    H.set_precision(prec)
    # FP32 mix chains: the fused run reads the packed FP32 copy of the Lanczos vector that its
    # residual launch writes, the callback run the FP64 vector
    assert H.mode == "mix" and bool(H.chain.layout(0) & 8) == (prec == 1)
    x0 = psi._storage.clone()
    E1, v1, k1, a1, b1 = lanczos_gnd_native(H, x0, 1e-12, 400, batch=8)
    n = x0.numel()

    def cb(_ctx, xp, yp, _s):
        _lib.lib().cxa_chain_matvec(H.chain.handle, xp, yp, _s)
        return 0

    f = _MATVEC_T(cb)
    fn = ctypes.cast(f, ctypes.c_void_p).value
    E2, v2, k2, a2, b2 = lanczos_gnd_native(None, x0, 1e-12, 400, batch=8, fn_ctx=(fn, None))
    H.set_precision(0)
    print("prec", prec, "fused", k1, E1, "two-launch", k2, E2, flush=True)
    assert k1 == k2
    np.testing.assert_allclose(a1, a2, rtol=1e-11, atol=1e-11 * np.abs(a2).max())
    np.testing.assert_allclose(b1, b2, rtol=1e-9, atol=1e-11 * np.abs(b2).max())
    assert abs(E1 - E2) <= 1e-11 * abs(E2)
    assert torch.linalg.vector_norm(v1 - v2).item() <= 1e-6


def test_lanczos_context_reuse_across_sizes():
    """Runs of different vector lengths share one context (cxa_lanczos_reshape within its
    allocation): a smaller operator after a larger one reuses it, and every run equals a run
    on a freshly made context bit for bit, graph replay included; a reshape past the
    allocation is refused."""
    import ctypes
    from cytnx_amd import _lib
    from cytnx_amd import lanczos as lz

    def fresh():
        torch.cuda.synchronize()
        for h in list(lz._CTX_CAP):
            _lib.lib().cxa_lanczos_destroy(h)
        lz._CTX_CACHE.clear()
        lz._CTX_CAP.clear()

    big = problem(64, seed=31)
    small = problem(56, seed=37)
    HA, HB = Heff2(*big), Heff2(*small)
    xa, xb = big[0]._storage, small[0]._storage
    assert xb.numel() < xa.numel()
    fresh()
    ref_b = lanczos_gnd_native(HB, xb, 1e-12, 120, batch=8)
    fresh()
    ref_a = lanczos_gnd_native(HA, xa, 1e-12, 120, batch=8)
    assert len(lz._CTX_CAP) == 1
    got_b = lanczos_gnd_native(HB, xb, 1e-12, 120, batch=8)
    assert len(lz._CTX_CAP) == 1  # the larger context, reshaped
    got_a = lanczos_gnd_native(HA, xa, 1e-12, 120, batch=8, graph=True)
    got_b2 = lanczos_gnd_native(HB, xb, 1e-12, 120, batch=8, graph=True)
    assert len(lz._CTX_CAP) == 1
    for got, ref in ((got_b, ref_b), (got_a, ref_a), (got_b2, ref_b)):
        assert got[0] == ref[0] and got[2] == ref[2]
        assert np.array_equal(got[3], ref[3]) and np.array_equal(got[4], ref[4])
        assert torch.equal(got[1], ref[1])
    ctx = next(iter(lz._CTX_CAP))
    cap_n, cap_m, _ = lz._CTX_CAP[ctx]
    assert _lib.lib().cxa_lanczos_reshape(ctx, cap_n + 1, cap_m) == _lib.ERR_INVALID
    assert _lib.lib().cxa_lanczos_reshape(ctx, cap_n, cap_m + 1) == _lib.ERR_INVALID
    fresh()


def test_chain_buffers_reused_from_the_pool_match_the_oracle():
    """Chains made one after another reuse the device blocks of dropped ones (the library keeps
    them in a pool, with their old contents): a second operator of the same shape, made after
    the first is dropped, still equals the oracle chain in FP64 (1e-12) and FP32 (5e-6), and so
    does a third after an FP32 Lanczos run of the second has filled its buffers."""
    import gc
    first = problem(96, seed=41)
    H = Heff2(*first)
    H.set_precision(1)
    H.matvec(first[0])
    lanczos_gnd_native(H, first[0]._storage, 1e-10, 40, batch=8)
    torch.cuda.synchronize()
    del H
    gc.collect()
    for seed in (43, 47):
        psi, L, M1, M2, R = problem(96, seed=seed)
        H = Heff2(psi, L, M1, M2, R)
        ref = oracle_heff2(psi, L, M1, M2, R).flat()
        assert relerr(H.matvec(psi)._storage.cpu().numpy(), ref) <= 1e-12
        H.set_precision(1)
        assert relerr(H.matvec(psi)._storage.cpu().numpy(), ref) <= 5e-6
        lanczos_gnd_native(H, psi._storage, 1e-10, 40, batch=8)
        torch.cuda.synchronize()
        del H
        gc.collect()


def test_lanczos_matches_oracle_recurrence_on_heff2():
    psi, L, M1, M2, R = problem(16, seed=17)
    H = Heff2(psi, L, M1, M2, R)
    E, v, k, a, b = lanczos_gnd_native(H, psi._storage, 1e-12, 300, batch=8)

    def mv(x):
        t = torch.from_numpy(x).cuda()
        y = torch.empty_like(t)
        H.apply_flat(t, y)
        return y.cpu().numpy()

    Eo, vo, ko = po.lanczos_gnd(mv, psi._storage.cpu().numpy(), 1e-12, 300)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)


def test_mixed_precision_refinement_within_1e10():
    psi, L, M1, M2, R = problem(64, seed=19)
    H = Heff2(psi, L, M1, M2, R)
    E64, v64, k64, _, _ = lanczos_gnd_native(H, psi._storage, 1e-12, 1000, batch=8, reorth=True)
    Emx, vmx, n1, n2 = ground_state_mixed(H, psi._storage, 1e-12, 1000, batch=8)
    print("E64", E64, "iters", k64, "mixed", Emx, "fp32 iters", n1, "fp64 iters", n2, flush=True)
    assert abs(Emx - E64) <= 1e-10 * max(1.0, abs(E64))
