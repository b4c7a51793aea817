"""GPU parity of the permute path (cxa_permute through Tensor.permute_ + contiguous_).

Bit-exact against the oracle's restated odometer loop on the permute sweep (rank-6 FP64 and
complex128 tensors, shapes (16,8,16,8,16,8) and (32,4,32,4,32,4), 32 seeded permutations),
on the reference's permute fixtures, and on edge cases (unit axes, rank 1, empty, every
element size, lazy composition)."""
import os

import numpy as np
import pytest
import torch

import pyoracle as po
from conftest import REF_FIXTURES
from cytnx_amd import fileio
from cytnx_amd.tensor import Tensor

pytestmark = pytest.mark.gpu

SWEEP_SHAPES = [(16, 8, 16, 8, 16, 8), (32, 4, 32, 4, 32, 4)]


def sweep_perms(n=32, rank=6, seed=20240):
    rng = np.random.default_rng(seed)
    return [rng.permutation(rank).tolist() for _ in range(n)]


def _bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a).view(np.uint8)


def _gpu_permute(a: np.ndarray, perm) -> np.ndarray:
    t = Tensor.from_numpy(a, device=0)
    t.permute_(perm)
    t.contiguous_()
    torch.cuda.synchronize()
    assert t.is_contiguous()
    return t.numpy()


# the sweep shapes, and ragged ones whose tiles end partway (short output runs, small axes next
# to the contiguous ones)
@pytest.mark.parametrize("shape", SWEEP_SHAPES + [(33, 4, 17, 4, 5, 4), (3, 64, 5, 2, 40, 2)])
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_permute_sweep_bit_exact(shape, dtype):
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, size=shape)
    if dtype == np.complex128:
        x = x + 1j * rng.uniform(-1, 1, size=shape)
    x = x.astype(dtype)
    for perm in sweep_perms():
        print("permute", shape, np.dtype(dtype).name, perm, flush=True)
        got = _gpu_permute(x, perm)
        ref = po.permute(x, perm)
        assert got.shape == ref.shape
        assert np.array_equal(_bits(got), _bits(ref)), perm


# batches of small transposes (the last two axes swapped, or two short axes moved past each
# other): tiles that batch several indices of the next outer axis, with batch counts that do not
# divide that axis (ragged last tiles)
BATCH_CASES = [((16, 8, 16, 8, 16, 8), [0, 1, 2, 3, 5, 4]), ((32, 4, 32, 4, 32, 4), [0, 1, 2, 3, 5, 4]),
               ((3, 16, 8), [0, 2, 1]), ((7, 4, 32), [0, 2, 1]), ((11, 8, 16), [0, 2, 1]),
               ((13, 3, 5), [0, 2, 1]), ((9, 5, 6, 3), [1, 0, 3, 2]), ((2, 3, 4, 5, 6), [0, 1, 2, 4, 3]),
               ((4, 100, 3), [0, 2, 1]), ((5, 3, 7, 2), [0, 1, 3, 2])]


@pytest.mark.parametrize("dtype", [np.bool_, np.int16, np.float32, np.float64, np.complex128])
def test_permute_batched_small_transposes_bit_exact(dtype):
    rng = np.random.default_rng(13)
    for shape, perm in BATCH_CASES:
        x = rng.uniform(-1, 1, size=shape)
        if dtype == np.complex128:
            x = x + 1j * rng.uniform(-1, 1, size=shape)
        if dtype == np.bool_:
            x = x > 0
        x = (x * 100).astype(dtype) if np.dtype(dtype).kind == "i" else x.astype(dtype)
        print("permute", shape, np.dtype(dtype).name, perm, flush=True)
        got = _gpu_permute(x, perm)
        assert np.array_equal(_bits(got), _bits(po.permute(x, perm)))


@pytest.mark.parametrize("dtype", [np.float32, np.complex64, np.int32, np.int16, np.uint64,
                                   np.bool_, np.float64])
def test_permute_all_element_sizes(dtype):
    rng = np.random.default_rng(11)
    for shape in [(3, 5, 7), (2, 1, 9, 1, 4), (130, 70), (64, 33, 3), (1, 1, 5), (6,)]:
        x = (rng.uniform(-50, 50, size=shape)).astype(dtype)
        for perm in [list(rng.permutation(len(shape))) for _ in range(4)]:
            print("permute", shape, np.dtype(dtype).name, perm, flush=True)
            assert np.array_equal(_gpu_permute(x, perm), np.transpose(x, perm))


@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.int16, np.bool_])
def test_identity_permute_streaming_copy(dtype):
    """Permutations that reduce to the identity (unit axes, axes that stay in order) are a
    streaming copy in 16-byte vectors, or a device copy when the bytes are not 16-byte whole."""
    rng = np.random.default_rng(17)
    for shape, perm in [((1 << 20,), [0]), ((3, 1, 1000), [1, 0, 2]), ((7, 5, 3), [0, 1, 2]),
                        ((1, 33, 1), [2, 1, 0]), ((5,), [0]), ((64, 64, 1, 3), [0, 2, 1, 3])]:
        x = rng.uniform(-1, 1, size=shape)
        if dtype == np.complex128:
            x = x + 1j * rng.uniform(-1, 1, size=shape)
        x = (x > 0) if dtype == np.bool_ else (x * 100).astype(dtype)
        print("identity permute", shape, np.dtype(dtype).name, perm, flush=True)
        got = _gpu_permute(x, perm)
        assert np.array_equal(_bits(got), _bits(np.ascontiguousarray(np.transpose(x, perm))))


def test_permute_empty_and_scalar_like():
    x = np.zeros((0, 4, 3))
    assert _gpu_permute(x, [2, 0, 1]).shape == (3, 0, 4)
    x = np.arange(5.0).reshape(1, 5, 1)
    assert np.array_equal(_gpu_permute(x, [2, 1, 0]), np.transpose(x, [2, 1, 0]))


def test_lazy_composition_moves_once():
    rng = np.random.default_rng(12)
    x = rng.standard_normal((4, 5, 6, 7))
    t = Tensor.from_numpy(x, device=0)
    t.permute_(3, 1, 0, 2).permute_(2, 0, 3, 1)
    ref = np.transpose(np.transpose(x, (3, 1, 0, 2)), (2, 0, 3, 1))
    t.contiguous_()
    assert np.array_equal(t.numpy(), ref)


def test_dense_permute_fixtures():
    u3 = fileio.read_unitensor(os.path.join(REF_FIXTURES, "denseutensor3.cytnx"))
    p2 = fileio.read_unitensor(os.path.join(REF_FIXTURES, "densepermu2.cytnx"))
    u4 = fileio.read_unitensor(os.path.join(REF_FIXTURES, "denseutensor4.cytnx"))
    p1 = fileio.read_unitensor(os.path.join(REF_FIXTURES, "densepermu1.cytnx"))
    got2 = _gpu_permute(u3.blocks[0].logical(), [1, 0])
    got1 = _gpu_permute(u4.blocks[0].logical(), [1, 0, 3, 2])
    assert np.array_equal(_bits(got2), _bits(p2.blocks[0].logical()))
    assert np.array_equal(_bits(got1), _bits(p1.blocks[0].logical()))


@pytest.mark.parametrize("case,perm", [(1, [1, 2, 0]), (2, [1, 0]), (3, [1, 0])])
def test_block_permute_fixtures_per_block(case, perm):
    T = po.from_saved(fileio.read_unitensor(os.path.join(REF_FIXTURES, "permute_T%d.cytnx" % case)))
    ans = po.from_saved(fileio.read_unitensor(os.path.join(REF_FIXTURES,
                                                           "permute_ans%d.cytnx" % case)))
    lookup = {tuple(t): b for t, b in zip(ans.itoi, ans.blocks)}
    for t, blk in zip(T.itoi, T.blocks):
        key = tuple(t[m] for m in perm)
        got = _gpu_permute(blk, perm)
        assert np.array_equal(_bits(got), _bits(lookup[key]))


@pytest.mark.parametrize("case,perm", [(1, [1, 2, 0]), (2, [1, 0]), (3, [1, 0])])
def test_unitensor_block_permute_fixtures(case, perm):
    """UniTensor.permute of the reference's block tensors (one batched block-permute plan per
    layout) against the reference's answers, bit for bit."""
    from cytnx_amd.unitensor import UniTensor
    T = UniTensor.Load(os.path.join(REF_FIXTURES, "permute_T%d.cytnx" % case), 0)
    ans = po.from_saved(fileio.read_unitensor(os.path.join(REF_FIXTURES,
                                                           "permute_ans%d.cytnx" % case)))
    P = T.permute(perm)
    lookup = {tuple(t): b for t, b in zip(ans.itoi, ans.blocks)}
    assert len(P.itoi()) == len(ans.itoi)
    for i, t in enumerate(P.itoi()):
        got = P.get_block_(i).cpu().numpy()
        assert np.array_equal(_bits(got), _bits(lookup[tuple(t)]))


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.complex64, np.complex128])
def test_block_permute_plan_small_and_large_blocks(dtype):
    """A block tensor whose blocks go through both routes of a block-permute plan: the small
    ones in one batched launch, the 72 x 64 x 64 one (> 2^15 elements) through cxa_permute;
    every block bit-exact against numpy.transpose."""
    from cytnx_amd.unitensor import BD_BRA, BD_KET, Bond, UniTensor
    rng = np.random.default_rng(5)
    bl = Bond(BD_KET, [[0], [1]], [72, 3])
    bs = Bond(BD_KET, [[0], [1]], [64, 2])
    br = Bond(BD_BRA, [[0], [1], [2]], [64, 5, 2])
    dt = {np.float32: 4, np.float64: 3, np.complex64: 2, np.complex128: 1}[dtype]
    T = UniTensor([bl, bs, br], ["l", "s", "r"], 2, dt, device=0)
    x = rng.standard_normal(T.storage().numel())
    if np.issubdtype(dtype, np.complexfloating):
        x = x + 1j * rng.standard_normal(x.size)
    T._storage = torch.from_numpy(x.astype(dtype)).to("cuda:0")
    sizes = [int(np.prod(T._shapes[i].astype(np.int64))) for i in range(T.Nblocks())]
    assert max(sizes) > 1 << 15 and min(sizes) < 1 << 15
    for perm in ([2, 0, 1], [1, 2, 0], [0, 2, 1], [0, 1, 2]):
        P = T.permute(perm)
        for i in range(T.Nblocks()):
            src = T.get_block_(i).cpu().numpy()
            j = P.itoi().index([T.itoi()[i][m] for m in perm])
            got = P.get_block_(j).cpu().numpy()
            assert np.array_equal(_bits(got), _bits(np.transpose(src, perm)))
