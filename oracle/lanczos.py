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


EPMACH = 2.0 ** -52   # machine_precision (block_lanczos.f90:927-943)


def _localro(q1, q2):
    """localro (:1430-1470): Q_{j+1} orthogonalised against the vectors of Q_j
    one after another, then its columns normalised."""
    q2 = q2.copy()
    for m in range(q1.shape[1]):
        q2 -= np.outer(q1[:, m], q2.T @ q1[:, m])
    return q2 / np.sqrt((q2 * q2).sum(axis=0))


def _lowerbound(om, j, k, eta):
    """lowerbound (:1848-1868).  Its loop ``do i=k-2,1`` runs only when
    k - 2 <= 1, so for k >= 4 the bound is k - 1."""
    if k in (1, 2):
        return 1
    lb = k - 1
    for i in range(k - 2, 2):              # do i=k-2,1 (step +1)
        if abs(om[j + 1, i]) > eta:
            lb = i
        else:
            break
    return lb


def _upperbound(om, j, k, eta):
    """upperbound (:1872-1890)."""
    ub = k + 1
    for i in range(k + 1, j):              # do i=k+1,j-1
        if abs(om[j + 1, i]) >= eta:
            ub = i
        else:
            break
    return ub


def _mpro(om, jj, b, L, q, eps, orthlim, l2nd, intervals):
    """mpro (:1710-1781) and mpro_mgs (:1785-1845) at cycle jj (1-based):
    Q_{j+1} (``q``, modified in place) orthogonalised vector by vector against
    the stored blocks of the intervals, its columns normalised; the
    intervals found are widened by one block each side and used again in the
    next cycle.  Returns (lro, l2nd, intervals, number of re-orthogonalisations)."""
    eta = EPMACH ** 0.75

    def mgs(ivs):
        for lo, hi in ivs:
            for k in range(lo, hi + 1):
                for m in range(b):
                    l = L[:, (k - 1) * b + m]
                    q[:] -= np.outer(l, q.T @ l)
        q[:] /= np.sqrt((q * q).sum(axis=0))

    if l2nd:
        mgs(intervals)
        for lo, hi in intervals:
            for l in range(lo, hi + 1):
                om[jj + 1, l] = om[l, jj + 1] = eps
        return True, False, [], 1
    ivs, k, lro = [], 1, False
    while True:                            # the goto-10 loop
        if om[jj + 1, k] >= orthlim:
            lro = True
            lo, hi = _lowerbound(om, jj, k, eta), _upperbound(om, jj, k, eta)
            ivs.append([lo, hi])
            k = k + hi
        else:
            k += 1
        if not k < jj - 1:
            break
    if not lro:
        return False, False, [], 0
    mgs(ivs)
    for iv in ivs:
        for l in range(iv[0], iv[1] + 1):
            om[jj + 1, l] = om[l, jj + 1] = eps
        if iv[0] > 1:
            iv[0] -= 1
        iv[1] += 1
    return True, True, ivs, 1


def block_lanczos(matvec, q0: np.ndarray, ncycles: int, orthotype: int = 0) -> dict:
    """``matvec(X)`` returns H X for an (n, b) block; q0 holds orthonormal columns.

    orthotype 1 adds partial re-orthogonalisation (pro, :1474-1526).  The
    omega recurrence (update_omega, :1530-1639) estimates the loss of
    orthogonality from the spectral norms of A_j and B_j and the smallest
    singular value of B_j.  When an estimate reaches sqrt(eps), pro_mgs
    (:1643-1706) orthogonalises the working Q_j and Q_{j+1} against the
    stored blocks 1..j-1 vector by vector, normalises Q_j and applies localro.
    Otherwise localro alone runs (:1192).  The stored Lanczos vectors (Ritz
    vectors, later re-orthogonalisations) are the blocks as they were at the
    start of their cycle; the recurrence uses the modified working Q_j.

    orthotype 2 is the modified partial re-orthogonalisation (mpro,
    :1710-1781): the same omega recurrence; Q_{j+1} alone is orthogonalised
    against the intervals of stored blocks around the estimates that reach
    sqrt(eps) (bounds where the estimates fall below eps^(3/4)), and against
    the widened intervals once more in the next cycle; no localro.

    Returns T, eigval (ascending), Y (eigenvectors of T), ritz (normalised
    L Y), L (the Lanczos vectors), Qnext / Bnext (the block and the triangle
    of the last cycle's QR, which T leaves out) and npro (re-orthogonalisations).
    """
    q0 = np.asarray(q0, dtype=np.float64)
    n, b = q0.shape
    nvec = b * ncycles
    L = np.zeros((n, nvec))
    T = np.zeros((nvec, nvec))
    q_prev = np.zeros((n, b))
    b_prev = np.zeros((b, b))
    q_cur = q0.copy()
    eps = EPMACH * b * np.sqrt(float(n))
    orthlim = np.sqrt(EPMACH)
    om = np.zeros((ncycles + 2, ncycles + 2))
    anorm = np.zeros(ncycles + 1)
    bnorm = np.zeros(ncycles + 1)
    amats, bmats = [], []
    npro = 0
    l2nd, intervals = False, []
    for j in range(ncycles):
        c0 = j * b
        L[:, c0:c0 + b] = q_cur
        u = matvec(q_cur) - q_prev @ b_prev.T
        a = q_cur.T @ u
        r = u - q_cur @ a
        q_next, b_next = np.linalg.qr(r)
        amats.append(a)
        bmats.append(b_next)
        q_prev, q_cur = q_cur.copy(), q_next
        jj = j + 1                                   # the reference's 1-based cycle
        lro = False
        if orthotype in (1, 2) and 1 < jj < ncycles:
            sv = lambda m: np.linalg.svd(m, compute_uv=False)
            if jj == 2:
                anorm[1] = sv(amats[0])[0]
                bnorm[1] = sv(bmats[0])[0]
                s2 = sv(bmats[1])
                bnorm[2] = s2[0]
                om[1, 1] = om[2, 2] = eps
                om[1, 2] = om[2, 1] = eps / s2[-1]
            anorm[jj] = sv(amats[jj - 1])[0]
            sb = sv(bmats[jj - 1])
            bnorm[jj] = sb[0]
            inv = 1.0 / sb[-1]
            om[jj, jj + 1] = om[jj + 1, jj] = om[jj + 1, jj + 1] = eps
            for k in range(1, jj):
                x = bnorm[k + 1] * om[jj, k + 1] + bnorm[jj] * om[jj - 1, k] + (anorm[jj] + anorm[k]) * om[jj, k]
                if k > 1:
                    x += bnorm[k] * om[jj, k - 1]
                om[jj + 1, k] = om[k, jj + 1] = inv * x
            if orthotype == 2:
                lro, l2nd, intervals, did = _mpro(om, jj, b, L, q_cur, eps, orthlim, l2nd, intervals)
                npro += did
            elif any(om[jj + 1, k] >= orthlim for k in range(1, jj)):
                lro = True
                npro += 1
                for k in range(1, jj):
                    for m in range(b):
                        l = L[:, (k - 1) * b + m]
                        d1 = q_prev.T @ l
                        d2 = q_cur.T @ l
                        q_prev -= np.outer(l, d1)
                        q_cur -= np.outer(l, d2)
                q_prev /= np.sqrt((q_prev * q_prev).sum(axis=0))
                q_cur = _localro(q_prev, q_cur)
                for k in range(1, jj + 1):
                    om[jj + 1, k] = eps
                    om[jj, k] = eps
        if orthotype == 1 and not lro:
            q_cur = _localro(q_prev, q_cur)
        T[c0:c0 + b, c0:c0 + b] = np.triu(a) + np.triu(a, 1).T
        if j + 1 < ncycles:
            T[c0 + b:c0 + 2 * b, c0:c0 + b] = b_next
            T[c0:c0 + b, c0 + b:c0 + 2 * b] = b_next.T
        b_prev = b_next
    lam, Y = np.linalg.eigh(T, UPLO="L")
    ritz = L @ Y
    ritz /= np.sqrt((ritz * ritz).sum(axis=0))
    return {"T": T, "eigval": lam, "Y": Y, "ritz": ritz, "L": L, "Qnext": q_cur, "Bnext": b_prev,
            "npro": npro}


def subspace_guess(hii, hsub, blocksize, kind=8, subdim=800, minc=1e-12):
    """init_vec_doublesubspace (kind 8, :443-744) / init_vec_singlesubspace
    (kind 9, :748-923).  A subspace is the subdim ISs of lowest (ascending)
    or highest (descending) H_jj, extended while the next H_jj is within
    1e-6 of the last; ``hsub(sel)`` returns H[sel][:, sel] (|h| <= minc
    dropped, the diagonal from hii).  Kind 8: columns 1..bs/2 are the lowest
    eigenvectors of the ascending subspace, columns bs/2+1..bs eigenvectors
    bs/2+1..bs (ascending eigenvalue order, as the reference's column loop
    takes them) of the descending one.  Kind 9: the lowest bs eigenvectors of
    the descending subspace."""
    hii = np.asarray(hii, dtype=np.float64)
    n = hii.size

    def block(descending):
        order = np.argsort(-hii if descending else hii, kind="stable")
        sd = min(subdim, n)
        while sd < n and abs(hii[order[sd - 1]] - hii[order[sd]]) < 1e-6:
            sd += 1
        sel = order[:sd]
        A = np.array(hsub(sel), dtype=np.float64)
        A[np.abs(A) <= minc] = 0.0
        A[np.arange(sd), np.arange(sd)] = hii[sel]
        return sel, np.linalg.eigh(A)[1]

    V = np.zeros((n, blocksize))
    if kind == 9:
        sel, U = block(True)
        V[sel, :] = U[:, :blocksize]
        return V
    h = blocksize // 2
    sel, U = block(False)
    V[sel, :h] = U[:, :h]
    sel, U = block(True)
    V[sel, h:] = U[:, h:blocksize]
    return V


def tranmom(ritz: np.ndarray, travec: np.ndarray) -> np.ndarray:
    """travec . r_n for every Ritz vector: the tm(ndim, vec, mtm) products that
    get_tranmom_1 (source/adclib/fspacetrial.f90:1076-1156, called at
    source/adc/adc2_opa.f90:766) forms over the records ritzvecs_incore writes."""
    return ritz.T @ np.asarray(travec, dtype=np.float64)
