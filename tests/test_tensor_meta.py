"""Host-side Tensor metadata semantics (no device needed): lazy permute composition, the
contiguity flag, reshape validation.  Mirrors tests/Tensor_test.cpp:293-330 of the reference."""
import numpy as np
import pytest

from cytnx_amd.tensor import Tensor


def test_permute_is_lazy_and_composes():
    t = Tensor([2, 3, 4])
    p = t.permute(1, 2, 0)
    assert p.shape() == [3, 4, 2]
    assert not p.is_contiguous()
    assert p.same_data(t)
    assert p.mapper() == [1, 2, 0]
    # mapper[j] is the storage axis of logical axis j; invmapper is its inverse
    assert [p.mapper()[p.invmapper()[i]] for i in range(3)] == [0, 1, 2]
    q = p.permute(2, 0, 1)  # back to identity: (1,2,0) composed with (2,0,1)
    assert q.shape() == [2, 3, 4]
    assert q.is_contiguous()


def test_permute_rejects_bad_arguments():
    t = Tensor([2, 3])
    with pytest.raises(ValueError):
        t.permute_(0, 0)
    with pytest.raises(ValueError):
        t.permute_(0, 1, 2)
    with pytest.raises(ValueError):
        t.permute_(0, 2)


def test_numpy_view_follows_mapper():
    a = np.arange(24.0).reshape(2, 3, 4)
    t = Tensor.from_numpy(a)
    t.permute_(2, 0, 1)
    assert np.array_equal(t.numpy(), np.transpose(a, (2, 0, 1)))


def test_reshape_validation():
    t = Tensor([2, 3, 4])
    with pytest.raises(ValueError):
        t.reshape_(5, -1)
    with pytest.raises(ValueError):
        t.reshape_(-1, -1)
    t.reshape_(6, -1)
    assert t.shape() == [6, 4]


def test_cpu_move_is_refused():
    t = Tensor([2, 3])
    t.permute_(1, 0)
    with pytest.raises(NotImplementedError):
        t.contiguous_()


def test_from_storage_materialises_lazy_conjugates():
    """The kernels read raw memory, so a torch conj / neg view must not reach them: torch's SVD
    returns Vh as a conjugate view, which .contiguous() keeps."""
    import torch
    x = torch.randn(3, 4, dtype=torch.complex128)
    t = Tensor.from_storage(x.conj().contiguous(), [3, 4])
    assert not t.storage().is_conj()
    assert torch.equal(t.storage().reshape(3, 4), torch.conj_physical(x))
    n = Tensor.from_storage(torch._neg_view(x), [3, 4])
    assert not n.storage().is_neg() and torch.equal(n.storage().reshape(3, 4), -x)
