"""Ground-state energy of the open spin-1/2 Heisenberg chain from its Bethe ansatz.

Test infrastructure: a known answer for DMRG on long chains, where exact diagonalisation is out
of reach.  H = sum_{i=1}^{L-1} S_i . S_{i+1} with free ends (the XXZ MPO at J = Jz = 1).  For
even L the ground state has N = L/2 magnons with real rapidities lambda_1 < ... < lambda_N > 0
solving the logarithmic Bethe equations of the open chain (F. C. Alcaraz, M. N. Barber,
M. T. Batchelor, R. J. Baxter, G. R. W. Quispel, J. Phys. A 20 (1987) 6397; E. K. Sklyanin,
J. Phys. A 21 (1988) 2375), with quantum numbers I_j = j:

    2 L arctan(2 lambda_j) = pi I_j + sum_{k != j} [arctan(lambda_j - lambda_k)
                                                     + arctan(lambda_j + lambda_k)],

and E = (L - 1)/4 - sum_j 2 / (4 lambda_j^2 + 1) (each magnon lowers the ferromagnetic energy
(L - 1)/4 by 2 / (4 lambda^2 + 1)).  tests/test_bethe_host.py checks the solver against dense
exact diagonalisation for L = 4 .. 14.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import root


def open_xxx_rapidities(L: int) -> np.ndarray:
    if L < 2 or L % 2:
        raise ValueError("even L >= 2")
    N = L // 2
    I = np.arange(1, N + 1, dtype=np.float64)

    def F(lam):
        d = np.arctan(lam[:, None] - lam[None, :]).sum(1)           # k = j term is 0
        s = np.arctan(lam[:, None] + lam[None, :]).sum(1) - np.arctan(2 * lam)  # k != j
        return 2 * L * np.arctan(2 * lam) - np.pi * I - d - s

    lam0 = 0.5 * np.tan(np.pi * I / (2 * L + 2))  # the free-magnon guess
    sol = root(F, lam0, method="hybr", tol=1e-14)
    lam = sol.x
    # polish with Newton steps on the analytic Jacobian
    for _ in range(5):
        step = np.linalg.solve(_jacobian(L, lam), F(lam))
        lam = lam - step
        if np.abs(step).max() <= 1e-16 * max(1.0, np.abs(lam).max()):
            break
    if np.abs(F(lam)).max() > 1e-12:
        raise RuntimeError("Bethe equations did not converge for L=%d" % L)
    return lam


def _jacobian(L: int, lam: np.ndarray) -> np.ndarray:
    """d F_j / d lambda_m of the logarithmic Bethe equations above."""
    dd = lam[:, None] - lam[None, :]
    ss = lam[:, None] + lam[None, :]
    gd = 1.0 / (1.0 + dd ** 2)
    gs = 1.0 / (1.0 + ss ** 2)
    J = gd - gs                                   # m != j: +1/(1+d^2) - 1/(1+s^2)
    diag = (4.0 * L / (1.0 + 4.0 * lam ** 2)
            - (gd.sum(1) - 1.0)                   # sum_{k != j} 1/(1+(l_j-l_k)^2)
            - (gs.sum(1) - gs.diagonal())         # sum_{k != j} 1/(1+(l_j+l_k)^2)
            )
    np.fill_diagonal(J, diag)
    return J


def open_xxx_ground_energy(L: int) -> float:
    lam = open_xxx_rapidities(L)
    return float((L - 1) / 4.0 - np.sum(2.0 / (4.0 * lam ** 2 + 1.0)))
