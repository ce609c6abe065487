"""Exact ADC(2)-s eigenvalues by Lowdin partitioning.  Test oracle only.

In ADC(2)-s the 2h2p-2h2p block is diagonal (source/adclib/adc_ph.f90:1090-1096),
so an eigenvalue w of H = [[H11, C], [C^T, diag(D)]] below min(D) is a root of

    f_i(w) = lambda_i( H11 + C diag(1/(w - D)) C^T ) - w = 0,

and f_i is strictly decreasing there (d/dw of 1/(w-D) is negative), so the i-th
root is unique and Newton's method with
f_i'(w) = -sum_J (C^T x_i)_J^2 / (w - D_J)^2 - 1 converges from below.

This pins the Davidson solvers (oracle and GPU) independently of any Krylov
method: it only needs n1 x n1 dense eigensolves.
"""
from __future__ import annotations

import numpy as np


def roots(H11: np.ndarray, C: np.ndarray, D: np.ndarray, nroots: int,
          tol: float = 1e-14, maxit: int = 60) -> np.ndarray:
    """The ``nroots`` lowest eigenvalues; raises ValueError if one of them is
    not below min(D).  ``count_below`` tells how many are."""
    dmin = D.min()
    w0 = np.linalg.eigvalsh(H11)[:nroots]
    out = np.empty(nroots)
    for i in range(nroots):
        w = min(w0[i], dmin - 1e-3)
        for _ in range(maxit):
            s = 1.0 / (w - D)
            M = H11 + (C * s) @ C.T
            lam, vec = np.linalg.eigh(M)
            x = vec[:, i]
            f = lam[i] - w
            y = C.T @ x
            fp = -np.sum(y * y * s * s) - 1.0
            step = -f / fp
            w_new = w + step
            if w_new >= dmin:
                w_new = 0.5 * (w + dmin)
            if abs(w_new - w) < tol * max(1.0, abs(w)):
                w = w_new
                break
            w = w_new
        s = 1.0 / (w - D)
        lam = np.linalg.eigvalsh(H11 + (C * s) @ C.T)
        if w >= dmin or abs(lam[i] - w) > 1e3 * tol * max(1.0, abs(w)):
            raise ValueError("root %d is not below min(D); partitioning does not apply" % i)
        out[i] = w
    return out


def count_below(H11, C, D, nmax):
    """Number of eigenvalues of H below min(D) (at most nmax): by interlacing,
    root i lies below min(D) iff lambda_i(M(w)) < w as w -> min(D)-."""
    w = D.min() - 1e-9
    s = 1.0 / (w - D)
    lam = np.linalg.eigvalsh(H11 + (C * s) @ C.T)
    return int(min(nmax, np.sum(lam < w)))
