"""Block Lanczos pseudospectrum.  Test oracle only.

Restates run_block_lanczos (source/eigen/block_lanczos.f90:947-1256) in its
default form, orthotype 0 (no re-orthogonalisation; defaults.f90:139):

    U   = H Q_j - Q_{j-1} B_{j-1}^T                      (:1112-1121, B_0 = 0)
    A_j = Q_j^T U                                        (:1123-1124)
    R   = U - Q_j A_j                                    (:1129-1132)
    R   = Q_{j+1} B_j, B_j upper triangular (dgeqrf, dorgqr; :1137-1166)

T (nvec = blocksize * ncycles) is block tridiagonal.  The diagonal block j
is A_j's upper triangle mirrored, which is what the fill order of :1194-1217
leaves.  The sub-diagonal block (j+1, j) is B_j for j < ncycles.  The
eigenpairs of T come from dsbev on the lower band (diagmat_banded,
:2006-2052).  The Ritz vectors are L Y, with each column divided by its
2-norm (ritzvecs_incore, :2237-2300).

numpy.linalg.qr calls the same LAPACK pair (dgeqrf, dorgqr), so Q_{j+1} and
B_j carry the reference's sign convention.

Pinned in tests/test_oracle_pins.py by properties that the mathematics fixes:
  * the three-term recurrence H L = L T + Q_{m+1} B_m E_m^T;
  * orthonormality of L;
  * exactness of block Gauss quadrature,
    sum_n (f.r_n)^2 theta_n^k = f^T H^k f for f in span(Q_1), k < 2m;
  * eig(T) = eig(H) when the Krylov space is the whole space.
"""
from __future__ import annotations

import numpy as np


def block_lanczos(matvec, q0: np.ndarray, ncycles: int) -> dict:
    """``matvec(X)`` returns H X for an (n, b) block; q0 holds orthonormal columns.

    Returns T, eigval (ascending), Y (eigenvectors of T), ritz (normalised
    L Y), L (the Lanczos vectors), and Qnext / Bnext (the block and the
    triangle of the last cycle's QR, which T leaves out).
    """
    q0 = np.asarray(q0, dtype=np.float64)
    n, b = q0.shape
    nvec = b * ncycles
    L = np.zeros((n, nvec))
    T = np.zeros((nvec, nvec))
    q_prev = np.zeros((n, b))
    b_prev = np.zeros((b, b))
    q_cur = q0.copy()
    for j in range(ncycles):
        c0 = j * b
        L[:, c0:c0 + b] = q_cur
        u = matvec(q_cur) - q_prev @ b_prev.T
        a = q_cur.T @ u
        r = u - q_cur @ a
        q_next, b_next = np.linalg.qr(r)
        T[c0:c0 + b, c0:c0 + b] = np.triu(a) + np.triu(a, 1).T
        if j + 1 < ncycles:
            T[c0 + b:c0 + 2 * b, c0:c0 + b] = b_next
            T[c0:c0 + b, c0 + b:c0 + 2 * b] = b_next.T
        q_prev, q_cur, b_prev = q_cur, q_next, b_next
    lam, Y = np.linalg.eigh(T, UPLO="L")
    ritz = L @ Y
    ritz /= np.sqrt((ritz * ritz).sum(axis=0))
    return {"T": T, "eigval": lam, "Y": Y, "ritz": ritz, "L": L, "Qnext": q_cur, "Bnext": b_prev}


def tranmom(ritz: np.ndarray, travec: np.ndarray) -> np.ndarray:
    """travec . r_n for every Ritz vector: the tm(ndim, vec, mtm) products that
    get_tranmom_1 (source/adclib/fspacetrial.f90:1076-1156, called at
    source/adc/adc2_opa.f90:766) forms over the records ritzvecs_incore writes."""
    return ritz.T @ np.asarray(travec, dtype=np.float64)
