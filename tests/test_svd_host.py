"""Host logic of the truncated SVD: the block cut rule (src/linalg/Svd_truncate.cpp:163-190 with
block_truncation_helpers.hpp) and the dense rule (memcpyTruncation.cpp:13-60), on hand-made
singular value lists whose answers follow from the rules' text; and diagonal UniTensor metadata.
"""
import numpy as np
import pytest

from cytnx_amd.svd import _cut, _cut_dense
from cytnx_amd.unitensor import BD_KET, Bond, UniTensor


def test_block_cut_keepdim_and_degeneracy():
    s = np.array([5.0, 4.0, 3.0, 3.0, 1.0, 0.5])
    Smin, dropped, Sall = _cut(s, 3, 0.0, 1)
    # the 3rd largest is 3.0; the degenerate partner at the cut is kept too
    assert Smin == 3.0 and dropped == 2
    assert list(Sall) == sorted(s)
    Smin, dropped, _ = _cut(s, 10, 0.0, 1)       # keepdim beyond the count: keep all
    assert Smin == 0.5 and dropped == 0


def test_block_cut_err_and_mindim():
    s = np.array([5.0, 1e-3, 1e-9, 1e-12])
    Smin, dropped, _ = _cut(s, 4, 1e-6, 1)
    assert Smin == 1e-3 and dropped == 2
    Smin, dropped, _ = _cut(s, 4, 1.0, 1)       # err above all but the largest: keep one
    assert Smin == 5.0 and dropped == 3
    Smin, dropped, _ = _cut(s, 4, 1.0, 3)       # mindim keeps three
    assert Smin == 1e-9 and dropped == 1


def test_dense_cut_rule():
    s = np.array([5.0, 4.0, 3.0, 3.0, 1.0, 0.5])
    assert _cut_dense(s, 3, 0.0, 1) == 3          # no degeneracy rule
    assert _cut_dense(s, 10, 0.0, 1) == 6
    assert _cut_dense(s, 6, 2.0, 1) == 4          # trailing values below err dropped
    assert _cut_dense(s, 6, 10.0, 1) == 1         # at least one kept
    assert _cut_dense(s, 6, 10.0, 2) == 2         # mindim


def test_diag_unitensor_metadata():
    b = Bond(BD_KET, [[-1], [1], [3]], [2, 3, 1])
    S = UniTensor([b, b.redirect()], ["_aux_L", "_aux_R"], 1, 3, -1, is_diag=True)
    assert S.is_diag() and S.Nblocks() == 3
    assert S.storage().numel() == 6
    assert S.itoi() == [[0, 0], [1, 1], [2, 2]]
    S._storage[:] = __import__("torch").arange(1, 7, dtype=__import__("torch").float64)
    D = S.to_dense()
    np.testing.assert_array_equal(D, np.diag(np.arange(1.0, 7.0)))
    assert list(S.get_block_(1).numpy()) == [3.0, 4.0, 5.0]
    P = S.permute([1, 0])
    assert P.labels() == ["_aux_R", "_aux_L"] and P.is_diag()
    T = S.Transpose()
    assert T.bonds()[0].type() == BD_KET and T.labels() == ["_aux_R", "_aux_L"]
    with pytest.raises(ValueError):
        UniTensor([b, b], ["a", "b"], 1, 3, -1, is_diag=True)
    with pytest.raises(ValueError):
        UniTensor([b], ["a"], 1, 3, -1, is_diag=True)


# ---- the per-block floor (min_blockdim) and the reference's truncation fixtures ----------

import os  # noqa: E402

from cytnx_amd import fileio  # noqa: E402
from cytnx_amd.svd import _cut_blockdim, _sector_matrices, _split_overload  # noqa: E402

TRUNC = os.path.join(os.path.dirname(__file__), "golden", "ref", "truncate")


def test_min_blockdim_exhausting_keepdim_drops_every_candidate():
    """Three sectors of three values, floor 2 each, keepdim 3: the floors keep six values, the
    reduced keepdim is -3, so the three third values are all dropped
    (tests/linalg_test/Svd_truncate_test.cpp ReturnErrWhenMinBlockdimExhaustsKeepdim)."""
    blocks = [np.array([3.0, 2.0, 0.3]), np.array([5.0, 1.0, 0.2]), np.array([4.0, 0.9, 0.7])]
    keep, Sall, dropped = _cut_blockdim(blocks, 3, [2, 2, 2], 0.0, 1)
    assert keep == [2, 2, 2] and dropped == 3
    assert list(Sall) == [0.2, 0.3, 0.7]


def test_min_blockdim_keeps_each_sector():
    """keepdim 1 alone keeps one sector; with floor 1 per sector all three survive
    (Svd_truncate_test.cpp MinBlockdimKeepsEachSector)."""
    blocks = [np.array([3.0, 2.0, 0.3]), np.array([5.0, 1.0, 0.2]), np.array([4.0, 0.9, 0.7])]
    Smin, _, _ = _cut(np.concatenate(blocks), 1, 0.0, 1)
    assert [int(np.count_nonzero(b >= Smin)) for b in blocks] == [0, 1, 0]
    keep, _, dropped = _cut_blockdim(blocks, 1, [1], 0.0, 1)   # one element: every block
    assert keep == [1, 1, 1] and dropped == 6


def test_min_blockdim_candidates_compete_for_the_rest():
    blocks = [np.array([3.0, 2.0, 0.3]), np.array([5.0, 1.0, 0.2]), np.array([4.0, 0.9, 0.7])]
    # floor 1 on the first block only: 2 of keepdim 3 left for the 8 candidates (5 and 4 win)
    keep, Sall, dropped = _cut_blockdim(blocks, 3, [1, 0, 0], 0.0, 1)
    assert keep == [1, 1, 1] and dropped == 6 and Sall.size == 8
    # a block whose floor covers it is kept whole and offers no candidates
    keep, Sall, dropped = _cut_blockdim(blocks, 4, [5, 0, 0], 0.0, 1)
    assert keep == [3, 1, 0] and Sall.size == 6 and dropped == 5
    # no candidates at all: every block whole, nothing dropped
    keep, Sall, dropped = _cut_blockdim(blocks, 2, [3, 3, 3], 0.0, 1)
    assert keep == [3, 3, 3] and Sall.size == 0 and dropped == 0
    # err drops candidates down to zero when every block has a floor (floor of the count 0)
    keep, Sall, dropped = _cut_blockdim(blocks, 9, [1, 1, 1], 10.0, 1)
    assert keep == [1, 1, 1] and dropped == 6
    # without any floor the err loop stops at one value
    keep, _, dropped = _cut_blockdim(blocks, 9, [0, 0, 0], 10.0, 1)
    assert keep == [0, 1, 0] and dropped == 8


def test_min_blockdim_rejects_bad_lengths():
    blocks = [np.array([1.0]), np.array([2.0])]
    with pytest.raises(ValueError):
        _cut_blockdim(blocks, 1, [], 0.0, 1)
    with pytest.raises(ValueError):
        _cut_blockdim(blocks, 1, [1, 1, 1], 0.0, 1)


def test_truncate_overload_parsing():
    assert _split_overload((1e-3, False), {}) == (None, (1e-3, False))
    assert _split_overload(([1, 0], 1e-3), {}) == ([1, 0], (1e-3,))
    assert _split_overload((), {"min_blockdim": (2,)}) == ([2], ())


def _sector_svals(path):
    T = UniTensor.Load(path)           # host copy: only the block grouping is used here
    return [(sec.key, np.linalg.svd(sec.m.numpy(), compute_uv=False))
            for sec in _sector_matrices(T)]


@pytest.mark.parametrize("dt", ["F64", "C128"])
def test_cut_rules_reproduce_the_reference_truncation_fixtures(dt):
    """keepdim 5 on the U(1) tensor with exponentially falling singular values, without and with
    min_blockdim {1, 1, 0, 2, 0}: the kept sectors, their sizes and values are those of the
    reference's answer files (tests/linalg_test/Gesvd_truncate_test.cpp U1ExpSvalsTest,
    U1ExpSvalsMinblockdimTest; the singular values here come from numpy's SVD of each sector)."""
    secs = _sector_svals(os.path.join(TRUNC, "src", "sym_UT_U1_exp_Svals_%s.cytnx" % dt))
    svals = [s for _, s in secs]
    Smin, _, _ = _cut(np.concatenate(svals), 5, 0.0, 1)
    plain = [int(np.count_nonzero(s >= Smin)) for s in svals]
    floor, _, _ = _cut_blockdim(svals, 5, [1, 1, 0, 2, 0], 0.0, 1)
    for keep, name in ((plain, ""), (floor, "_minblockdim")):
        ans = fileio.read_unitensor(os.path.join(TRUNC, "ans",
                                                 "sym_UT_U1_exp_Svals_%s%s.cytnx" % (dt, name)))
        live = [i for i, k in enumerate(keep) if k]
        assert [[secs[i][0]] for i in live] == ans.bonds[0].qnums
        assert [keep[i] for i in live] == ans.bonds[0].degs
        for i, blk in zip(live, ans.blocks):
            want = blk.logical().real
            np.testing.assert_allclose(svals[i][:keep[i]], want, rtol=1e-9, atol=1e-15)
