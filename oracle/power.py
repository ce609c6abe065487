"""Power method for the highest eigenpair.  Test oracle only.

Restates power_method (source/eigen/power.f90:22-100).  The guess
(power_guess, :307-463) is the highest eigenvector of the stored matrix on
the 800 ISs of highest H_jj.  Those ISs are sorted descending, and the count
grows while the next H_jj is within 1e-6 of the last; off-diagonal elements
|h| <= 1e-12 are dropped as the stored matrix drops them.  The iterations
(power_iterations, :467-524) are, up to maxit = 350:
    v2 = H v1,  E = v1.v2,  r = |E v1 - v2|,  v2 /= |v2|,
stopping when r <= 1e-8, else v1 = v2.

Pinned in tests/test_oracle_pins.py by the dense spectrum: the converged E is
the largest eigenvalue.
"""
from __future__ import annotations

import numpy as np


def guess(hii, hsub, subdim=800, minc=1e-12):
    hii = np.asarray(hii, dtype=np.float64)
    n = hii.size
    order = np.argsort(-hii, kind="stable")
    sd = min(subdim, n)
    while sd < n and abs(hii[order[sd - 1]] - hii[order[sd]]) < 1e-6:
        sd += 1
    sel = order[:sd]
    A = np.array(hsub(sel), dtype=np.float64)
    A[np.abs(A) <= minc] = 0.0
    A[np.arange(sd), np.arange(sd)] = hii[sel]
    _, U = np.linalg.eigh(A)
    v = np.zeros(n)
    v[sel] = U[:, -1]
    return v


def power_method(matvec, v1, maxit=350, thrsh=1e-8):
    """Returns (E, v, iterations, converged)."""
    v1 = np.asarray(v1, dtype=np.float64).copy()
    e = 0.0
    for it in range(1, maxit + 1):
        v2 = matvec(v1)
        e = v1 @ v2
        r = np.sqrt(((e * v1 - v2) ** 2).sum())
        v2 = v2 / np.sqrt(v2 @ v2)
        if r <= thrsh:
            return e, v1, it, True
        v1 = v2
    return e, v1, maxit, False
