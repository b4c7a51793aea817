"""Host check of the orthonormalisation behind the DMRG starting state
(cytnx_amd.dmrg._orthonormal_rows): every matrix of a zero-padded batch comes back with
orthonormal rows spanning its own row space (the polar factor), padding kept zero, including a
square matrix (slowest Newton-Schulz convergence) and a nearly rank-deficient one (QR
fallback).  Reference: numpy's SVD, whose U V^T is the polar factor."""
import numpy as np
import torch

from cytnx_amd.dmrg import _orthonormal_rows


def test_orthonormal_rows_is_the_polar_factor():
    rng = np.random.default_rng(0)
    sizes = [(5, 9), (16, 32), (30, 31), (1, 2), (20, 40), (48, 48)]
    mp = np_ = 48
    Mb = torch.zeros((len(sizes), mp, np_), dtype=torch.float64)
    rows = torch.zeros((len(sizes), mp), dtype=torch.float64)
    for b, (m, n) in enumerate(sizes):
        Mb[b, :m, :n] = torch.from_numpy(rng.standard_normal((m, n)))
        rows[b, :m] = 1
    X = _orthonormal_rows(Mb, rows).numpy()
    for b, (m, n) in enumerate(sizes):
        x = X[b]
        assert not np.any(x[m:]) and not np.any(x[:, n:])
        U, _, Vt = np.linalg.svd(Mb[b, :m, :n].numpy(), full_matrices=False)
        assert np.abs(x[:m, :n] - U @ Vt).max() <= 1e-12
        assert np.abs(x[:m, :n] @ x[:m, :n].T - np.eye(m)).max() <= 1e-13


def test_orthonormal_rows_fallback_near_rank_deficiency():
    rng = np.random.default_rng(1)
    M = rng.standard_normal((8, 8))
    M[7] = M[6] + 1e-13 * rng.standard_normal(8)
    X = _orthonormal_rows(torch.from_numpy(M)[None], torch.ones((1, 8), dtype=torch.float64),
                          maxit=40).numpy()[0]
    assert np.abs(X @ X.T - np.eye(8)).max() <= 1e-13
