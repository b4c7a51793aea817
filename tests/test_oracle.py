"""Pins the CPU oracle before anything is checked against it.

Each oracle function is checked against something it does not share code with: numpy closed
forms, the reference's own fixture files (tests/golden/ref, copied from the reference's
tests/test_data_base), brute force on small inputs, and the reference's hard-coded known answers.
"""
import itertools
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import REF_FIXTURES
from cytnx_amd import fileio


def _ref(name):
    return os.path.join(REF_FIXTURES, name)


# ---------------------------------------------------------------------------- permute (a2)
@pytest.mark.parametrize("dtype", [np.float64, np.complex128, np.float32, np.int32])
def test_permute_matches_transpose(dtype):
    rng = np.random.default_rng(1)
    for rank in range(1, 7):
        shape = tuple(int(x) for x in rng.integers(1, 5, size=rank))
        a = rng.standard_normal(shape)
        if np.issubdtype(dtype, np.complexfloating):
            a = a + 1j * rng.standard_normal(shape)
        a = a.astype(dtype)
        for perm in list(itertools.permutations(range(rank)))[:24]:
            assert np.array_equal(po.permute(a, perm), np.transpose(a, perm))


def test_permute_non_involutive_mapper_direction():
    # (1,2,0) and its inverse (2,0,1) differ: mixing mapper and invmapper fails here
    a = np.arange(2 * 3 * 5, dtype=np.float64).reshape(2, 3, 5)
    out = po.permute(a, [1, 2, 0])
    assert out.shape == (3, 5, 2)
    assert out[2, 4, 1] == a[1, 2, 4]


def test_dense_permute_fixtures():
    u3 = fileio.read_unitensor(_ref("denseutensor3.cytnx")).blocks[0].logical()
    p2 = fileio.read_unitensor(_ref("densepermu2.cytnx")).blocks[0].logical()
    u4 = fileio.read_unitensor(_ref("denseutensor4.cytnx")).blocks[0].logical()
    p1 = fileio.read_unitensor(_ref("densepermu1.cytnx")).blocks[0].logical()
    assert np.array_equal(po.permute(u3, [1, 0]), p2)
    assert np.array_equal(po.permute(u4, [1, 0, 3, 2]), p1)


@pytest.mark.parametrize("case,perm", [(1, [1, 2, 0]), (2, [1, 0]), (3, [1, 0])])
def test_block_permute_fixtures(case, perm):
    # tests/BlockUniTensor_test.cpp:308-361 compares element values and existence
    t = po.from_saved(fileio.read_unitensor(_ref("permute_T%d.cytnx" % case)))
    ans = po.from_saved(fileio.read_unitensor(_ref("permute_ans%d.cytnx" % case)))
    p = t.permute(perm)
    assert np.array_equal(p.exists_mask(), ans.exists_mask())
    assert np.array_equal(p.to_dense(), ans.to_dense())


# ---------------------------------------------------------------------------- matmul (a5)
@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
@pytest.mark.parametrize("mkn", [(1, 1, 1), (5, 7, 3), (33, 17, 65), (4, 0, 3)])
def test_matmul(dtype, mkn):
    rng = np.random.default_rng(2)
    M, K, N = mkn
    A = rng.standard_normal((M, K)).astype(dtype)
    B = rng.standard_normal((K, N)).astype(dtype)
    if dtype == np.complex128:
        A = A + 1j * rng.standard_normal((M, K))
        B = B + 1j * rng.standard_normal((K, N))
    np.testing.assert_allclose(po.matmul(A, B), A @ B, rtol=1e-13, atol=1e-13)


# ---------------------------------------------------------------------------- tensordot (a4)
def test_tensordot_closed_form():
    rng = np.random.default_rng(3)
    L = rng.standard_normal((3, 4, 5, 6))
    R = rng.standard_normal((5, 2, 3))
    out = po.tensordot(L, R, [2, 0], [0, 2])
    np.testing.assert_allclose(out, np.tensordot(L, R, axes=([2, 0], [0, 2])), rtol=1e-13,
                               atol=1e-13)


@pytest.mark.parametrize("case,labels2", [(1, "a aa bb cc"), (2, "a b bb cc"), (3, "a b c cc")])
def test_dense_contract_fixtures(case, labels2):
    # tests/DenseUniTensor_test.cpp:2594-2619: ut1 labels a,b,c,d contracted with ut2
    u1 = fileio.read_unitensor(_ref("denseutensor1.cytnx")).blocks[0].logical().astype(complex)
    u2 = fileio.read_unitensor(_ref("denseutensor2.cytnx")).blocks[0].logical().astype(complex)
    ans = fileio.read_unitensor(_ref("densecontres%d.cytnx" % case)).blocks[0].logical()
    la, lb = "a b c d".split(), labels2.split()
    comm = sorted(set(la) & set(lb))  # vec_intersect_ sorts the common labels
    out = po.tensordot(u1, u2, [la.index(c) for c in comm], [lb.index(c) for c in comm])
    np.testing.assert_allclose(out, ans, rtol=1e-12, atol=1e-12)


# ---------------------------------------------------------------------------- block contract (a9)
@pytest.mark.parametrize("case,la,lb", [(1, "a b", "b c"), (2, "a b", "b c"),
                                        (3, "a b c", "c d e")])
def test_block_contract_fixtures(case, la, lb):
    # tests/BlockUniTensor_test.cpp:679-716
    L = po.from_saved(fileio.read_unitensor(_ref("contract_L%d.cytnx" % case)))
    R = po.from_saved(fileio.read_unitensor(_ref("contract_R%d.cytnx" % case)))
    ans = po.from_saved(fileio.read_unitensor(_ref("contract_ans%d.cytnx" % case)))
    L.labels, R.labels = la.split(), lb.split()
    C = po.block_contract(L, R)
    assert C.itoi == ans.itoi
    for c, a in zip(C.blocks, ans.blocks):
        assert c.shape == a.shape
        np.testing.assert_allclose(c, a, rtol=1e-12, atol=1e-12)


def _rand_block(bonds, labels, rr, rng, dtype=np.float64):
    T = po.OBlockTensor(bonds, labels, rr, dtype)
    for i, b in enumerate(T.blocks):
        x = rng.standard_normal(b.shape)
        if dtype == np.complex128:
            x = x + 1j * rng.standard_normal(b.shape)
        T.blocks[i] = x.astype(dtype)
    return T


@pytest.mark.parametrize("dtype", [np.float64, np.complex128])
def test_block_contract_dense_closed_form(dtype):
    rng = np.random.default_rng(4)
    K = po.KET
    bl = po.OBond(K, [-1, 0, 2], [2, 3, 1])
    bp = po.OBond(K, [1, -1], [1, 1])
    bw = po.OBond(K, [0, 2, -2], [3, 1, 1])
    # A(l, s, r) with r = l + s ; B(w, r, r') ; contract over r (label "r")
    br = po.OBond(po.BRA, [-2, -1, 0, 1, 3], [1, 2, 3, 2, 1])
    A = _rand_block([bl, bp, br], ["l", "s", "r"], 2, rng, dtype)
    B = _rand_block([bw, br.redirect(), po.OBond(po.BRA, [-2, 0, 2, 1], [2, 2, 1, 3])],
                    ["w", "r", "x"], 1, rng, dtype)
    C = po.block_contract(A, B)
    dense = np.tensordot(A.to_dense(), B.to_dense(), axes=([2], [1]))
    np.testing.assert_allclose(C.to_dense(), dense, rtol=1e-12, atol=1e-12)
    # blocks absent from the zero-flux list are exactly zero in the dense product
    assert np.all(dense[~C.exists_mask()] == 0)
    # output labels: left free then right free; rowrank arithmetic of :964-968
    assert C.labels == ["l", "s", "w", "x"]
    assert C.rowrank == 2 + 1 - 0 - 0


def test_block_full_contract_scalar():
    rng = np.random.default_rng(5)
    b1 = po.OBond(po.KET, [0, 1, -1], [2, 1, 3])
    A = _rand_block([b1, b1.redirect()], ["a", "b"], 1, rng)
    B = _rand_block([b1.redirect(), b1], ["b", "a"], 1, rng)
    s = po.block_contract(A, B)
    expect = np.sum(A.to_dense() * B.to_dense().T)
    assert abs(s - expect) < 1e-12


# ---------------------------------------------------------------------------- tridiag (a15)
@pytest.mark.parametrize("n", [1, 2, 3, 10, 57, 120])
def test_tridiag_eigh(n):
    rng = np.random.default_rng(6 + n)
    d = rng.standard_normal(n)
    e = rng.standard_normal(max(0, n - 1))
    w, Z, rc = po.tridiag_eigh(d, e)
    assert rc == 0
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    np.testing.assert_allclose(w, np.linalg.eigvalsh(T), rtol=0, atol=1e-12 * max(1, abs(w).max()))
    # rows are orthonormal eigenvectors
    np.testing.assert_allclose(Z @ Z.T, np.eye(n), atol=1e-12)
    np.testing.assert_allclose(T @ Z.T, Z.T * w, atol=1e-11 * max(1, abs(w).max()))


def test_tridiag_degenerate_and_split():
    d = np.array([1.0, 1.0, 1.0, 2.0, 2.0])
    e = np.array([0.0, 0.0, 1e-300, 0.0])
    w, Z, rc = po.tridiag_eigh(d, e)
    assert rc == 0
    np.testing.assert_allclose(w, [1, 1, 1, 2, 2], atol=1e-15)


# ---------------------------------------------------------------------------- Lanczos (a14)
def test_lanczos_known_27x27():
    # tests/linalg_test/Lanczos_Gnd_test.cpp:377-390: A = arange(27*27) + transpose,
    # Tin = arange(27), eigenvalue -1628.9964650426593 within 1e-5
    A = np.arange(27 * 27, dtype=np.float64).reshape(27, 27)
    A = A + A.T
    E, v, na = po.lanczos_gnd(lambda x: A @ x, np.arange(27.0), CvgCrit=1e-14, Maxiter=10000)
    assert abs(E - (-1628.9964650426593)) < 1e-5
    assert abs(E - np.linalg.eigvalsh(A)[0]) < 1e-9


def test_lanczos_block_operator_fixture():
    # tests/linalg_test/Lanczos_Gnd_test.cpp:392-423: block-diagonal 27x27 operator from
    # lan_block_{A,B,C}; eigenvalue within 1e-12 of Eigh, residual < 1e-6
    blocks = [fileio.read_tensor(_ref("lan_block_%s.cytn" % n)).logical() for n in "ABC"]
    H = np.zeros((27, 27))
    for i, b in enumerate(blocks):
        H[9 * i:9 * i + 9, 9 * i:9 * i + 9] = b
    E0 = min(np.linalg.eigvalsh(b)[0] for b in blocks)
    rng = np.random.default_rng(7)
    E, v, na = po.lanczos_gnd(lambda x: H @ x, rng.standard_normal(27), CvgCrit=1e-15,
                              Maxiter=10000)
    assert abs(E - E0) < 1e-12
    assert np.linalg.norm(H @ v - E * v) < 1e-6


def test_lanczos_stops_on_cvgcrit():
    rng = np.random.default_rng(8)
    X = rng.standard_normal((60, 60))
    H = X + X.T
    E, v, na = po.lanczos_gnd(lambda x: H @ x, rng.standard_normal(60), CvgCrit=1e9,
                              Maxiter=50)
    # |dE| < huge CvgCrit stops after the second matvec (two tridiagonal entries)
    assert na == 2
