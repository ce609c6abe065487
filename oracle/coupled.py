"""The coupled (Schur-complement) Davidson preconditioner, restated in numpy.
Test oracle only.

No counterpart in the reference, whose correction steps (dpr / olsen,
source/eigen/block_davidson.f90:1288-1419) divide the residual by
theta - H_jj element by element.  With H = [[H11, C], [C^T, diag(D)]] (the
ADC(2)-s matrix: the 2h2p block is diagonal, source/adclib/adc_ph.f90:1090-1096)
the coupled step solves

    M(theta) t = r,   M(theta) = [[H11 - theta, C], [C^T, diag(D) - sigma]]

through the 1h1p Schur complement A0(sigma) - theta, A0 = H11 - C P C^T,
P = diag(1 / (D - sigma)):

    t1 = (A0 - theta)^-1 (r1 - C P r2),   t2 = P (r2 - C^T t1).

Pinned by: the step solves M(theta) t = r (test_oracle_coupled); the number of
eigenvalues of A0(sigma) below sigma equals the number of eigenvalues of H
below sigma for sigma < min D (Sylvester's law of inertia on the block LDL^T
factorisation), checked against a dense eigen-decomposition of H.
"""
from __future__ import annotations

import numpy as np

from . import dfsigma


def coupling_matrix(df: dfsigma.DF) -> np.ndarray:
    """C (n1 x n2) of the density-fitted restatement, column by column of C^T."""
    n1 = df.n1
    Ct = np.empty((df.D2.size, n1))
    for j in range(n1):
        e = np.zeros(n1)
        e[j] = 1.0
        Ct[:, j] = df.couple_up(e)
    return Ct.T.copy()


def schur_complement(H11: np.ndarray, C: np.ndarray, D: np.ndarray, sigma: float) -> np.ndarray:
    """A0 = H11 - C diag(1/(D - sigma)) C^T."""
    if not sigma < D.min():
        raise ValueError("sigma must lie below every 2h2p diagonal element")
    return H11 - (C / (D - sigma)) @ C.T


def coupled_step(H11, C, D, sigma, theta, R):
    """t = M(theta)^-1 r for the columns of R (ndim x nb), one theta per column."""
    n1 = H11.shape[0]
    A0 = schur_complement(H11, C, D, sigma)
    P = 1.0 / (D - sigma)
    T = np.empty_like(R)
    for i in range(R.shape[1]):
        r1, r2 = R[:n1, i], R[n1:, i]
        t1 = np.linalg.solve(A0 - theta[i] * np.eye(n1), r1 - C @ (P * r2))
        T[:n1, i] = t1
        T[n1:, i] = P * (r2 - C.T @ t1)
    return T


def count_below(A0: np.ndarray, sigma: float) -> int:
    """Eigenvalues of H below sigma, from the inertia of A0 - sigma."""
    return int(np.sum(np.linalg.eigvalsh(A0) < sigma))
