"""The .cytnx / .cytn writer against the reference's own files (tests/golden/ref, written by
Cytnx's Save in the unversioned format: UniTensor magic 555, Tensor magic 888).

* Each fixture loaded into a UniTensor / Tensor and written back in that format reproduces the
  reference's file byte for byte: header flags, bond records, block enumeration order (itoi),
  tensor records and storages.
* The versioned format (556 / 889 plus version words, the one the reference writes today) reads
  back to the same content, and differs from the legacy image only by those words.
* UniTensor.Save / Tensor.Save: the extension rule, diagonal and dense tensors, names.
"""
import glob
import os
import struct

import numpy as np
import pytest
import torch

from conftest import REF_FIXTURES
from cytnx_amd import fileio
from cytnx_amd.tensor import Tensor
from cytnx_amd.unitensor import BD_KET, Bond, UniTensor

UTENS = sorted(glob.glob(os.path.join(REF_FIXTURES, "*.cytnx")) +
               glob.glob(os.path.join(REF_FIXTURES, "truncate", "*", "*.cytnx")))
TENS = sorted(glob.glob(os.path.join(REF_FIXTURES, "*.cytn")))


def _first_diff(a: bytes, b: bytes) -> int:
    for i in range(min(len(a), len(b))):
        if a[i] != b[i]:
            return i
    return -1 if len(a) == len(b) else min(len(a), len(b))


def _same_content(a, b):
    assert (a.uten_type, a.is_braket_form, a.is_tag, a.is_diag, a.rowrank, a.name, a.labels) == \
        (b.uten_type, b.is_braket_form, b.is_tag, b.is_diag, b.rowrank, b.name, b.labels)
    assert a.bonds == b.bonds
    assert a.itoi == b.itoi
    assert len(a.blocks) == len(b.blocks)
    for x, y in zip(a.blocks, b.blocks):
        assert x.dtype == y.dtype
        np.testing.assert_array_equal(x.logical(), y.logical())


@pytest.mark.parametrize("path", UTENS, ids=[os.path.basename(p) for p in UTENS])
def test_unitensor_fixture_bytes(path):
    with open(path, "rb") as f:
        orig = f.read()
    assert struct.unpack_from("<I", orig)[0] == 555
    U = UniTensor.Load(path)
    mine = fileio.unitensor_bytes(U.to_saved(), legacy=True)
    assert len(mine) == len(orig) and _first_diff(mine, orig) == -1, \
        "first differing byte %d" % _first_diff(mine, orig)


@pytest.mark.parametrize("path", UTENS, ids=[os.path.basename(p) for p in UTENS])
def test_unitensor_versioned_round_trip(path, tmp_path):
    src = fileio.read_unitensor(path)
    out = tmp_path / "x.cytnx"
    UniTensor.Load(path).Save(str(out))
    mine = out.read_bytes()
    assert struct.unpack_from("<II", mine) == (556, 1)
    _same_content(src, fileio.read_unitensor(str(out)))
    # the versioned image is the legacy one plus one version word per UniTensor / Tensor header
    ntens = 1 if src.uten_type == fileio.UTEN_DENSE else len(src.blocks)
    assert len(mine) == os.path.getsize(path) + 4 * (1 + ntens)


@pytest.mark.parametrize("path", TENS, ids=[os.path.basename(p) for p in TENS])
def test_tensor_fixture_bytes(path, tmp_path):
    src = fileio.read_tensor(path)
    assert src.contiguous
    T = Tensor.Load(path)
    with open(path, "rb") as f:
        assert fileio.tensor_bytes(T.numpy(), legacy=True) == f.read()
    T.Save(str(tmp_path / "t"))
    back = fileio.read_tensor(str(tmp_path / "t.cytn"))
    assert back.dtype == src.dtype
    np.testing.assert_array_equal(back.logical(), src.logical())


def test_lazily_permuted_tensor_is_written_contiguous(tmp_path):
    a = np.arange(24.0).reshape(2, 3, 4)
    T = Tensor.from_numpy(a).permute_(2, 0, 1)
    assert not T.is_contiguous()
    T.Save(str(tmp_path / "p.cytn"))
    s = fileio.read_tensor(str(tmp_path / "p.cytn"))
    assert s.contiguous and s.shape == [4, 2, 3]
    np.testing.assert_array_equal(s.logical(), np.transpose(a, (2, 0, 1)))


def test_block_unitensor_save_load(tmp_path):
    b = Bond(BD_KET, [-1, 1, 3], [2, 3, 1])
    p = Bond(BD_KET, [0, 2], [1, 1])
    T = UniTensor([b, p.redirect(), b.redirect()], ["a", "s", "c"], 2, 3, -1, name="A0")
    assert T.Nblocks() > 0
    rng = np.random.default_rng(3)
    for i in range(T.Nblocks()):
        T.put_block(rng.standard_normal(T.get_block_(i).shape), i)
    T.Save(str(tmp_path / "blk"))  # extension added to a bare name
    assert os.path.exists(tmp_path / "blk.cytnx")
    R = UniTensor.Load(str(tmp_path / "blk.cytnx"))
    assert R.labels() == T.labels() and R.rowrank() == T.rowrank() and R._name == "A0"
    assert [x.qnums() for x in R.bonds()] == [x.qnums() for x in T.bonds()]
    np.testing.assert_array_equal(R.to_dense(), T.to_dense())
    saved = fileio.read_unitensor(str(tmp_path / "blk.cytnx"))
    assert saved.is_braket_form is False and saved.is_tag is True  # (KET, BRA, BRA), rowrank 2


def test_diag_and_dense_save(tmp_path):
    b = Bond(BD_KET, [-2, 0, 2], [1, 2, 1])
    S = UniTensor([b, b.redirect()], ["_aux_L", "_aux_R"], 1, 3, -1, is_diag=True)
    vals = [np.array([4.0]), np.array([3.0, 2.0]), np.array([1.0])]
    for i, v in enumerate(vals):
        S.put_block(v, i)
    S.Save(str(tmp_path / "s.cytnx"))
    saved = fileio.read_unitensor(str(tmp_path / "s.cytnx"))
    assert saved.is_diag and saved.is_braket_form
    assert [blk.logical().tolist() for blk in saved.blocks] == [[4.0], [3.0, 2.0], [1.0]]
    D = UniTensor([Bond(3), Bond(4)], ["i", "j"], 1, 3, -1)
    D._block = Tensor.from_numpy(np.arange(12.0).reshape(3, 4))
    D.Save(str(tmp_path / "d.cytnx"))
    R = UniTensor.Load(str(tmp_path / "d.cytnx"))
    np.testing.assert_array_equal(R.to_dense(), np.arange(12.0).reshape(3, 4))
    t = Tensor.from_numpy(np.arange(6, dtype=np.complex128).reshape(2, 3))
    t.Save(str(tmp_path / "t"))
    np.testing.assert_array_equal(Tensor.Load(str(tmp_path / "t.cytn")).numpy(),
                                  np.arange(6, dtype=np.complex128).reshape(2, 3))
    assert torch.is_tensor(Tensor.Load(str(tmp_path / "t.cytn")).storage())
