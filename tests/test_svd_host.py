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
