"""GPU parity of the contraction path: Matmul, Gemm_Batch, Tensordot and block-sparse
UniTensor contraction, against the CPU oracle and the reference's fixtures.

FP64 / complex128 tolerance: normwise relative error <= 1e-12 (the north-star bar for
contractions); FP32 / complex64: <= 2e-6."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import REF_FIXTURES
from cytnx_amd import fileio, linalg
from cytnx_amd.tensor import Tensor
from cytnx_amd.unitensor import BD_BRA, BD_KET, Bond, Contract, UniTensor

pytestmark = pytest.mark.gpu

TOL = {np.float64: 1e-12, np.complex128: 1e-12, np.float32: 2e-6, np.complex64: 2e-6}


def relerr(got, ref):
    ref = np.asarray(ref)
    d = np.linalg.norm((np.asarray(got) - ref).ravel())
    n = np.linalg.norm(ref.ravel())
    return d / n if n > 0 else d


def rand(rng, shape, dtype):
    x = rng.standard_normal(shape)
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * rng.standard_normal(shape)
    return x.astype(dtype)


def ref_dtype(dtype):
    return np.complex128 if np.issubdtype(dtype, np.complexfloating) else np.float64


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.complex64])
def test_matmul(dtype):
    rng = np.random.default_rng(21)
    for M, K, N in [(1, 1, 1), (15, 16, 17), (16, 16, 16), (33, 7, 65), (100, 257, 31),
                    (246, 82, 82), (5, 0, 4), (64, 300, 1)]:
        print("matmul", np.dtype(dtype).name, M, K, N, flush=True)
        A, B = rand(rng, (M, K), dtype), rand(rng, (K, N), dtype)
        out = linalg.Matmul(Tensor.from_numpy(A, 0), Tensor.from_numpy(B, 0))
        rd = ref_dtype(dtype)
        ref = po.matmul(A.astype(rd), B.astype(rd))
        assert relerr(out.numpy(), ref) <= TOL[dtype], (M, K, N)


@pytest.mark.parametrize("dtype", [np.float64, np.float32, np.complex128, np.complex64])
def test_gemm_batch_groups_alpha_beta(dtype):
    """Two groups with their own alpha / beta (complex Scalars for complex dtypes, as
    Gemm_Batchfunc_oii passes them), against alpha A.B + beta C from the oracle's matmul."""
    rng = np.random.default_rng(22)
    shapes = [(3, 4, 5), (17, 33, 9), (40, 2, 70), (1, 1, 1), (64, 80, 48)]
    cplx = np.issubdtype(dtype, np.complexfloating)
    alpha = [1.0 - 0.25j, -0.5 + 2.0j] if cplx else [1.0, -0.5]
    beta = [0.0, 2.0 - 1.5j] if cplx else [0.0, 2.0]
    group = [2, 3]
    a, b, c, ref = [], [], [], []
    rd = ref_dtype(dtype)
    for i, (M, K, N) in enumerate(shapes):
        g = 0 if i < 2 else 1
        A, B, C = rand(rng, (M, K), dtype), rand(rng, (K, N), dtype), rand(rng, (M, N), dtype)
        a.append(Tensor.from_numpy(A, 0))
        b.append(Tensor.from_numpy(B, 0))
        c.append(Tensor.from_numpy(C, 0))
        ref.append(alpha[g] * po.matmul(A.astype(rd), B.astype(rd)) + beta[g] * C.astype(rd))
    linalg.Gemm_Batch(alpha, a, b, beta, c, group)
    for i, (got, r) in enumerate(zip(c, ref)):
        err = relerr(got.numpy(), r)
        print("gemm_batch", np.dtype(dtype).name, shapes[i], "err %.2e" % err, flush=True)
        assert err <= TOL[dtype]


def test_gemm_batch_rejects_complex_scalars_for_real_operands():
    A = Tensor.from_numpy(np.ones((2, 2)), 0)
    with pytest.raises(ValueError):
        linalg.Gemm_Batch([1j], [A], [A], [0.0], [Tensor.from_numpy(np.ones((2, 2)), 0)], [1])


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_tensordot_fused_and_copied_operands(dtype):
    rng = np.random.default_rng(23)
    cases = [((3, 4, 5, 6), (5, 2, 3), [2, 0], [0, 2]),   # L free (1,3): copy path
             ((6, 7, 8), (8, 9), [2], [0]),               # plain matrices
             ((4, 5, 6), (6, 5, 3), [1, 2], [1, 0]),      # K order differs between sides
             ((2, 3, 4, 5), (2, 3, 7), [0, 1], [0, 1])]   # K leading in L: strided view
    for sl, sr, il, ir in cases:
        print("tensordot", sl, sr, il, ir, flush=True)
        L, R = rand(rng, sl, dtype), rand(rng, sr, dtype)
        out = linalg.Tensordot(Tensor.from_numpy(L, 0), Tensor.from_numpy(R, 0), il, ir)
        assert relerr(out.numpy(), po.tensordot(L, R, il, ir)) <= TOL[dtype]


@pytest.mark.parametrize("case,labels2", [(1, "a aa bb cc"), (2, "a b bb cc"), (3, "a b c cc")])
def test_dense_contract_fixtures(case, labels2):
    u1 = UniTensor.Load(os.path.join(REF_FIXTURES, "denseutensor1.cytnx"), 0).astype(1)
    u2 = UniTensor.Load(os.path.join(REF_FIXTURES, "denseutensor2.cytnx"), 0).astype(1)
    ans = fileio.read_unitensor(os.path.join(REF_FIXTURES, "densecontres%d.cytnx" % case))
    u1.set_labels("a b c d".split())
    u2.set_labels(labels2.split())
    out = Contract(u1, u2)
    assert relerr(out.to_dense(), ans.blocks[0].logical()) <= 1e-12


@pytest.mark.parametrize("case,la,lb", [(1, "a b", "b c"), (2, "a b", "b c"),
                                        (3, "a b c", "c d e")])
def test_block_contract_fixtures(case, la, lb):
    L = UniTensor.Load(os.path.join(REF_FIXTURES, "contract_L%d.cytnx" % case), 0)
    R = UniTensor.Load(os.path.join(REF_FIXTURES, "contract_R%d.cytnx" % case), 0)
    ans = po.from_saved(fileio.read_unitensor(os.path.join(REF_FIXTURES,
                                                           "contract_ans%d.cytnx" % case)))
    L.set_labels(la.split())
    R.set_labels(lb.split())
    out = Contract(L, R)
    assert out.itoi() == ans.itoi
    for i, a in enumerate(ans.blocks):
        got = out.get_block_(i).cpu().numpy()
        assert got.shape == a.shape
        assert relerr(got, a) <= 1e-12 or np.abs(a).max() == 0


def _rand_ut(bonds, labels, rr, rng, dtype=np.float64):
    ut = UniTensor(bonds, labels, rr, 3 if dtype == np.float64 else 1, device=0)
    x = rand(rng, (ut.storage().numel(),), dtype)
    ut._storage = torch.from_numpy(x).to("cuda:0")
    return ut


def _oracle_of(ut):
    blocks = [ut.get_block_(i).cpu().numpy() for i in range(ut.Nblocks())]
    bonds = [po.OBond(b.type(), b.qnums(), b.getDegeneracies()) for b in ut.bonds()]
    return po.OBlockTensor(bonds, ut.labels(), ut.rowrank(), itoi=ut.itoi(), blocks=blocks)


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("seed", range(4))
def test_block_contract_random_vs_oracle(dtype, seed):
    rng = np.random.default_rng(100 + seed)
    bl = Bond(BD_KET, [[-1], [0], [2], [1]], [2, 3, 1, 17])
    bp = Bond(BD_KET, [[1], [-1]], [2, 1])  # a non-unit leg, so some permutes need copies
    br = Bond(BD_BRA, [[-2], [-1], [0], [1], [3], [2]], [1, 20, 3, 2, 1, 33])
    bw = Bond(BD_KET, [[0], [2], [-2]], [3, 1, 1])
    A = _rand_ut([bl, bp, br], ["l", "s", "r"], 2, rng, dtype)
    B = _rand_ut([bw, br.redirect(), Bond(BD_BRA, [[-2], [0], [2], [1], [3]], [2, 18, 1, 3, 5])],
                 ["w", "r", "x"], 1, rng, dtype)
    for perm in ([0, 1, 2], [2, 0, 1], [1, 2, 0]):
        Ap = A.permute(perm)  # moves data; exercises in-place and copied operand reads
        print("block contract", np.dtype(dtype).name, seed, perm, flush=True)
        C = Contract(Ap, B)
        ref = po.block_contract(_oracle_of(Ap), _oracle_of(B))
        assert C.itoi() == ref.itoi
        assert C.labels() == ref.labels and C.rowrank() == ref.rowrank
        assert relerr(C.to_dense(), ref.to_dense()) <= 1e-12


def test_block_full_contraction_scalar_and_kron():
    rng = np.random.default_rng(31)
    b1 = Bond(BD_KET, [[0], [1], [-1]], [2, 1, 3])
    A = _rand_ut([b1, b1.redirect()], ["a", "b"], 1, rng)
    B = _rand_ut([b1, b1.redirect()], ["b", "a"], 1, rng)  # KET b meets BRA b, BRA a meets KET a
    s = Contract(A, B)
    assert s.rank() == 0
    ref = po.block_contract(_oracle_of(A), _oracle_of(B))
    assert abs(s.item() - ref) <= 1e-12 * max(1.0, abs(ref))
    # no common label: every output block is an outer product
    B2 = B.relabel(["c", "d"])
    K = Contract(A, B2)
    dense = np.multiply.outer(A.to_dense(), B2.to_dense())
    assert relerr(K.to_dense(), dense) <= 1e-12


def test_contract_rejects_bra_bra():
    b1 = Bond(BD_KET, [[0], [1]], [1, 1])
    A = UniTensor([b1, b1.redirect()], ["a", "b"], device=0)
    B = UniTensor([b1, b1.redirect()], ["c", "b"], device=0)  # BRA meets BRA on "b"
    with pytest.raises(ValueError):
        Contract(A, B)


@pytest.mark.parametrize("kind,split", [(-1, 0), (-1, 1), (5, -1), (0, 1), (1, 1), (2, 1),
                                        (3, 1), (4, 0), (5, 1)])
def test_block_contract_forced_kernel_variants(kind, split):
    """Every tile shape, the row-streaming path and workgroup split-K give the oracle's answer
    (FP64 1e-12), including on blocks whose N exceeds the row path's limit."""
    from cytnx_amd import _lib
    from cytnx_amd.device import stream_handle
    from cytnx_amd.unitensor import contract_meta, plan_contraction
    rng = np.random.default_rng(200 + kind + 10 * split)
    bl = Bond(BD_KET, [[-1], [0], [2], [1]], [2, 3, 1, 37])
    bw = Bond(BD_KET, [[0], [2], [-2]], [3, 1, 1])
    br = Bond(BD_BRA, [[-2], [-1], [0], [1], [3], [2]], [1, 20, 3, 2, 1, 33])
    A = _rand_ut([bl, bw.redirect(), br], ["l", "w", "r"], 2, rng)
    B = _rand_ut([br.redirect(), Bond(BD_BRA, [[-2], [0], [2], [1], [3]], [2, 70, 1, 3, 5])],
                 ["r", "x"], 1, rng)
    C, c1, c2 = contract_meta(A, B)
    plan = plan_contraction(A, B, C, c1, c2, 3)
    L = _lib.lib()
    try:
        _lib.check(L.cxa_block_contract_set_tile(plan.handle, kind))
        _lib.check(L.cxa_block_contract_set_split(plan.handle, split))
        out = torch.full((int(np.sum(np.prod(C._shapes.astype(np.int64), axis=1))),), np.nan,
                         dtype=torch.float64, device="cuda:0")
        _lib.check(L.cxa_block_contract_exec(plan.handle, A._storage.data_ptr(),
                                             B._storage.data_ptr(), out.data_ptr(),
                                             plan.workspace(), 0, stream_handle(0)))
        C._storage = out
        ref = po.block_contract(_oracle_of(A), _oracle_of(B))
        assert relerr(C.to_dense(), ref.to_dense()) <= 1e-12
        assert not torch.isnan(out).any()
    finally:
        _lib.check(L.cxa_block_contract_set_tile(plan.handle, -1))
        _lib.check(L.cxa_block_contract_set_split(plan.handle, -1))


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_block_contract_batched_and_large_operand_copies(dtype):
    """Operands whose contracted legs are not adjacent are permuted into the workspace: blocks
    up to 2^15 elements by one batched launch per operand, larger ones one by one (here A's
    80 x 64 x 64 block); both routes against the oracle."""
    rng = np.random.default_rng(77)
    bl = Bond(BD_KET, [[0], [1]], [80, 3])
    bs = Bond(BD_KET, [[0], [1]], [64, 2])
    br = Bond(BD_BRA, [[0], [1], [2]], [64, 5, 2])
    bx = Bond(BD_BRA, [[-1], [0], [1], [2]], [2, 3, 4, 1])
    A = _rand_ut([bl, bs, br], ["l", "s", "r"], 2, rng, dtype)
    B = _rand_ut([br.redirect(), bx, bl.redirect()], ["r", "x", "l"], 1, rng, dtype)
    assert max(int(np.prod(A._shapes[i].astype(np.int64))) for i in range(A.Nblocks())) > 1 << 15
    C = Contract(A, B)
    ref = po.block_contract(_oracle_of(A), _oracle_of(B))
    assert C.itoi() == ref.itoi
    assert relerr(C.to_dense(), ref.to_dense()) <= 1e-12


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_strided_batch_plan_vs_numpy(dtype):
    """cxa_strided_batch_plan through the block-contract exec: a batch over a leg carried
    through unchanged, operands read and written in place with interleaved strides."""
    from cytnx_amd import _lib
    from cytnx_amd.device import stream_handle
    from cytnx_amd.heff import strided_batch_plan
    rng = np.random.default_rng(9)
    lp, w, s, r, wp, sp = 7, 5, 2, 6, 4, 3
    X = rand(rng, (lp, w, s, r), dtype)                 # batch over r, K = (w, s)
    W = rand(rng, (w, s, wp, sp), dtype)
    dt = 3 if dtype == np.float64 else 1
    plan = strided_batch_plan(r, lp, wp * sp, w * s, (1, w * s * r, r), (0, wp * sp, 1),
                              (wp * sp, r * wp * sp), dtype=dt)
    Xd, Wd = torch.from_numpy(X.ravel()).cuda(), torch.from_numpy(W.ravel()).cuda()
    Y = torch.full((lp * r * wp * sp,), float("nan"), dtype=Xd.dtype, device="cuda")
    _lib.check(_lib.lib().cxa_block_contract_exec(plan.handle, Xd.data_ptr(), Wd.data_ptr(),
                                                  Y.data_ptr(), None, 0, stream_handle(0)))
    want = np.einsum("awsr,wsvt->arvt", X, W)           # (lp, r, wp, sp)
    assert relerr(Y.cpu().numpy().reshape(want.shape), want) <= 1e-12
    assert plan.flops == 2.0 * r * lp * wp * sp * w * s * (4 if dt == 1 else 1)
