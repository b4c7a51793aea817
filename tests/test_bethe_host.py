"""The Bethe-ansatz ground energy of the open Heisenberg chain (tests/bethe.py) against dense
exact diagonalisation of the same Hamiltonian in the Sz = 0 sector, L = 4 .. 14; the L = 100
value is what the full DMRG run (the L=100, chi=2048 workload) is checked against."""
import numpy as np
import pytest

from bethe import _jacobian, open_xxx_ground_energy, open_xxx_rapidities


def _ed(L):
    states = [s for s in range(1 << L) if bin(s).count("1") == L // 2]
    idx = {s: i for i, s in enumerate(states)}
    H = np.zeros((len(states), len(states)))
    for s in states:
        i = idx[s]
        for b in range(L - 1):
            u, v = (s >> b) & 1, (s >> (b + 1)) & 1
            H[i, i] += 0.25 if u == v else -0.25
            if u != v:
                H[idx[s ^ (1 << b) ^ (1 << (b + 1))], i] += 0.5
    return np.linalg.eigvalsh(H)[0]


@pytest.mark.parametrize("L", [2, 4, 6, 8, 10, 12, 14])
def test_bethe_matches_exact_diagonalisation(L):
    assert abs(open_xxx_ground_energy(L) - _ed(L)) <= 1e-12 * L


def test_bethe_jacobian_matches_finite_differences():
    L = 12
    lam = open_xxx_rapidities(L) * 1.01
    N = L // 2
    I = np.arange(1, N + 1)

    def F(x):
        d = np.arctan(x[:, None] - x[None, :]).sum(1)
        s = np.arctan(x[:, None] + x[None, :]).sum(1) - np.arctan(2 * x)
        return 2 * L * np.arctan(2 * x) - np.pi * I - d - s

    h = 1e-6
    fd = np.stack([(F(lam + h * e) - F(lam - h * e)) / (2 * h) for e in np.eye(N)], axis=1)
    assert np.abs(fd - _jacobian(L, lam)).max() <= 1e-7


def test_bethe_L100_value_and_bulk_limit():
    E = open_xxx_ground_energy(100)
    # per site it approaches the Hulthen value 1/4 - ln 2 from above (boundary cost > 0)
    assert 0.25 - np.log(2) < E / 100 < 0.25 - np.log(2) + 0.01
    assert abs(E - (-44.127739893)) < 1e-8
