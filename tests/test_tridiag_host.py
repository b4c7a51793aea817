"""The library's host tridiagonal ground-state solver (Sturm bisection + inverse iteration)
against numpy.linalg.eigh and against the oracle's implicit-QR solver (a different algorithm)."""
import ctypes

import numpy as np
import pytest

import pyoracle as po
from cytnx_amd import _lib


def ground(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    E = ctypes.c_double(0)
    z = np.zeros(a.size)
    rc = _lib.lib().cxa_tridiag_ground(a.size, a.ctypes.data_as(_lib.dblp),
                                       b.ctypes.data_as(_lib.dblp) if b.size else None,
                                       ctypes.byref(E), z.ctypes.data_as(_lib.dblp))
    _lib.check(rc)
    return E.value, z


@pytest.mark.parametrize("n", [1, 2, 3, 7, 30, 100, 300])
def test_ground_pair_vs_eigh_and_oracle(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) * 3
    b = rng.standard_normal(n - 1)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    w, V = np.linalg.eigh(T)
    E, z = ground(a, b)
    scale = max(1.0, np.abs(w).max())
    assert abs(E - w[0]) <= 1e-13 * scale
    assert abs(abs(z @ V[:, 0]) - 1) <= 1e-10
    assert np.linalg.norm(T @ z - E * z) <= 1e-12 * scale
    wo, Zo, rc = po.tridiag_eigh(a, b)
    assert rc == 0 and abs(wo[0] - E) <= 1e-13 * scale


def test_lanczos_like_spectrum_near_degenerate():
    # Lanczos tridiagonals often carry ghost copies of the lowest eigenvalue
    n = 60
    a = np.full(n, -1.0)
    a[::7] = -5.0
    b = np.full(n - 1, 1e-9)
    E, z = ground(a, b)
    w = np.linalg.eigvalsh(np.diag(a) + np.diag(b, 1) + np.diag(b, -1))
    assert abs(E - w[0]) <= 1e-13 * 5


def test_rejects_bad_input():
    with pytest.raises(_lib.CxaError):
        ground(np.array([1.0, np.nan]), np.array([1.0]))


def eigh_all(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    n = a.size
    w = np.zeros(n)
    z = np.zeros((n, n))
    rc = _lib.lib().cxa_tridiag_eigh(n, a.ctypes.data_as(_lib.dblp),
                                     b.ctypes.data_as(_lib.dblp) if b.size else None,
                                     w.ctypes.data_as(_lib.dblp), z.ctypes.data_as(_lib.dblp))
    _lib.check(rc)
    return w, z


@pytest.mark.parametrize("n", [1, 2, 5, 40, 120])
def test_all_pairs_vs_eigh(n):
    """cxa_tridiag_eigh (Tridiag with is_V): every pair against numpy.linalg.eigh."""
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal(n) * 2
    b = rng.standard_normal(n - 1)
    T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
    w_ref = np.linalg.eigvalsh(T)
    w, z = eigh_all(a, b)
    scale = max(1.0, np.abs(w_ref).max())
    assert np.allclose(w, w_ref, atol=1e-13 * scale, rtol=0)
    assert np.allclose(z @ z.T, np.eye(n), atol=1e-12)
    assert np.linalg.norm(T @ z.T - z.T * w) <= 1e-12 * scale * n


def test_all_pairs_degenerate_and_decoupled():
    # a zero coupling splits the matrix; repeated eigenvalues keep orthonormal vectors
    a = np.array([1.0, 1.0, 2.0, 1.0])
    b = np.array([0.0, 0.0, 0.0])
    w, z = eigh_all(a, b)
    assert np.allclose(w, [1, 1, 1, 2])
    assert np.allclose(z @ z.T, np.eye(4), atol=1e-14)
