"""Block Davidson restated from source/eigen/block_davidson.f90.  Test oracle only.

Follows the reference step by step (1-based Fortran ranges in the comments):

* set-up ``davinitialise`` (:149-242): blocksize default max(n+5, 2n),
  maxvec default 4*blocksize, rounded down to a multiple of the blocksize;
* guess ``initvec_ondiag`` (:463-532): unit vectors on the ``blocksize``
  smallest |H_jj| (ascending sort; ties broken by index here, the reference's
  heapsort ``dsortindxa1`` (source/utilities/misc.f90:492-591) leaves tie
  order unspecified); ``subdiag_guess`` restates ``initvec_subdiag``
  (:293-459); a given block stands for ``initvec_adc1`` (:246-289);
* loop ``run_block_davidson`` (:690-774): sigma of all ``currdim`` columns
  (``hxvec_incore`` :844-883, nmult += currdim), Rayleigh matrix (:954-967),
  ``dsyev`` (:971-1008), Ritz vectors (:1012-1025), residuals
  r = theta x - W y and norms (:1029-1085), convergence count (reset each
  iteration without deflation, :1077-1081);
* ``subspace_expansion`` (:1089-1143): ``deflate`` (:1147-1222), collapse rule,
  ``dpr`` (:1288-1342) / ``olsen`` (:1346-1419), ``subspace_projection``
  (:1226-1284), QR of the whole subspace ``qrortho`` (:1423-1465);
* ``wreigenpairs`` (:1469-1532): deflated roots copied back, sorted ascending.

Pinned by: dense ``eigh`` and the Lowdin-partitioned roots of ``schur`` at
small sizes (test_oracle_pins); the iteration trajectory itself (iteration and
nmult counts) depends on LAPACK sign choices and is not pinned beyond the
reference's own algebra.
"""
from __future__ import annotations

import numpy as np


class NotConverged(RuntimeError):
    pass


DEN_MIN = 1e-10


def guard(den):
    """Preconditioner denominators with |den| < DEN_MIN are replaced by
    +-DEN_MIN.  The reference divides unguarded (block_davidson.f90:1327,1384);
    with unit-vector guesses on 2h2p configurations the first Ritz values
    equal H_jj exactly and the unguarded step is 0/0."""
    den = np.asarray(den, dtype=np.float64)
    small = np.abs(den) < DEN_MIN
    return np.where(small, np.where(den < 0, -DEN_MIN, DEN_MIN), den)


def subdiag_guess(hii, hsub, blocksize, subdim=800, minc=1e-12):
    """initvec_subdiag (source/eigen/block_davidson.f90:293-459).

    The subdim ISs of lowest H_jj (ascending), extended while the next H_jj
    is within 1e-6 of the last.  ``hsub(sel)`` returns H[sel][:, sel].  The
    diagonal is taken from hii, and off-diagonal elements |h| <= minc are
    dropped, as the stored matrix drops them.  The lowest ``blocksize``
    eigenvectors of that block are scattered into an (ndim, blocksize) array."""
    hii = np.asarray(hii, dtype=np.float64)
    n = hii.size
    order = np.argsort(hii, kind="stable")
    sd = min(subdim, n)
    while sd < n and abs(hii[order[sd - 1]] - hii[order[sd]]) < 1e-6:
        sd += 1
    sel = order[:sd]
    A = np.array(hsub(sel), dtype=np.float64)
    A[np.abs(A) <= minc] = 0.0
    A[np.arange(sd), np.arange(sd)] = hii[sel]
    _, U = np.linalg.eigh(A)
    V = np.zeros((n, blocksize))
    V[sel, :] = U[:, :blocksize]
    return V


def davidson(sigma, hii, nstates, blocksize=0, maxvec=-1, tol=1e-7, maxiter=100,
             precon=2, deflate=True, guess=None, log=None):
    """Return dict(eigval, eigvec, niter, nmult, table)."""
    ndim = hii.size
    bs = blocksize if blocksize > 0 else max(nstates + 5, 2 * nstates)
    mv = maxvec if maxvec >= 0 else 4 * bs
    mv = int(np.floor(mv / bs)) * bs
    V = np.zeros((ndim, mv))
    if guess is None:
        idx = np.argsort(np.abs(hii), kind="stable")[:bs]
        V[idx, np.arange(bs)] = 1.0
    else:
        V[:, :guess.shape[1]] = guess
    W = np.zeros_like(V)
    ritz = np.zeros((ndim, bs))
    res = np.zeros((ndim, bs))
    reigval = np.zeros(bs)
    norm = np.zeros(bs)
    currdim = bs
    bs_curr = bs
    mv_curr = mv
    ns_curr = nstates
    nconv = 0
    nconv_prev = 0
    nmult = 0
    table = []
    converged = False
    for it in range(1, maxiter + 1):
        # hxvec_incore: all currdim columns
        W[:, :currdim] = sigma(V[:, :currdim])
        nmult += currdim
        R = V[:, :currdim].T @ W[:, :currdim]
        val, vec = np.linalg.eigh(R)
        Y = vec[:, :bs_curr]
        reigval[:bs_curr] = val[:bs_curr]
        ritz[:, :bs_curr] = V[:, :currdim] @ Y
        res[:, :bs_curr] = -(W[:, :currdim] @ Y) + reigval[:bs_curr] * ritz[:, :bs_curr]
        norm[:bs_curr] = np.sqrt(np.sum(res[:, :bs_curr] ** 2, axis=0))
        if not deflate:
            nconv = 0
        nconv += int(np.sum(norm[:ns_curr] < tol))
        table.append((it, reigval[:ns_curr].copy(), norm[:ns_curr].copy()))
        if log is not None:
            log(it, reigval[:ns_curr], norm[:ns_curr])
        if nconv == nstates:
            converged = True
            break
        # ---- subspace_expansion ----
        newly = deflate and nconv > nconv_prev
        if newly:
            convmap = [i for i in range(bs_curr) if norm[i] < tol]
            for i in range(1, nconv - nconv_prev + 1):
                indx = bs + 1 - nconv_prev - i - 1          # 0-based column
                src = convmap[i - 1]
                ritz[:, [indx, src]] = ritz[:, [src, indx]]
                reigval[[indx, src]] = reigval[[src, indx]]
                res[:, [indx, src]] = res[:, [src, indx]]
            d = nconv - nconv_prev
            bs_curr -= d
            mv_curr -= d
            ns_curr -= d
            currdim -= d
        collapse = not (currdim <= mv_curr - bs_curr)
        if newly:
            collapse = True
        if collapse:
            V[:] = 0.0
            V[:, :bs_curr] = ritz[:, :bs_curr]
            ilast = bs_curr
        else:
            ilast = currdim
        for i in range(bs_curr):
            if precon == 1:
                V[:, ilast + i] = res[:, i] / guard(reigval[i] - hii)
            else:
                cd = 1.0 / guard(hii - reigval[i])
                xz = ritz[:, i] @ (res[:, i] * cd)
                xy = ritz[:, i] @ (ritz[:, i] * cd)
                alpha = xz / xy
                V[:, ilast + i] = alpha * cd * ritz[:, i] - cd * res[:, i]
        n = 2 * bs_curr if collapse else currdim + bs_curr
        if deflate and nconv > 0:
            conv = ritz[:, bs - nconv:bs]
            ov = conv.T @ V[:, :n]
            V[:, :n] -= conv @ ov
        q, _ = np.linalg.qr(V[:, :n])
        V[:, :n] = q
        currdim = n
        nconv_prev = nconv
    if not converged:
        raise NotConverged("Not all vectors have converged...")
    # wreigenpairs
    if deflate and nconv_prev > 0:
        reigval[ns_curr:nstates] = reigval[bs - nconv_prev:bs]
        ritz[:, ns_curr:nstates] = ritz[:, bs - nconv_prev:bs]
    order = np.argsort(reigval[:nstates], kind="stable")
    return dict(eigval=reigval[:nstates][order].copy(), eigvec=ritz[:, :nstates][:, order].copy(),
                niter=it, nmult=nmult, table=table)
