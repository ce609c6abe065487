"""Imaginary-time relaxation with the XSIL (Lanczos-Liu) integrator.  Test oracle only.

Restates xsil_rlx (source/propagation/relaxation.f90:139-309) with its helpers.
States s = 1..nstates are relaxed one after another from the initial vectors.
Each timestep of a state does the following.

  * collapse_subspace (:1163-1203): when the variational subspace is full,
    drop the oldest timestep's krydim + 1 vectors and its rows/columns.
  * get_lancvecs_current (:1207-1326): the first vector is psi,
    orthogonalised against the converged states 1..s-1 one after another and
    normalised.  Then krydim Lanczos steps follow
        r = H q - beta_{j-1} v,  alpha_j = q.r,  r -= alpha_j q,  beta_j = |r|
    and the (krydim+1)-th vector r / beta_krydim, which only enters the
    matrix elements.  Within the block H is tridiagonal (alpha, beta) and S = I.
  * subspace_matrices (:1330-1381): between a vector i of an earlier timestep
    and a vector j of the current one,
        S_ij = l_i.l_j,
        H_ij = beta_{k-1} l_i.l_{j-1} + alpha_k l_i.l_j + beta_k l_i.l_{j+1},
    with k the position of j in its block and no beta_0 term.
  * solve_subspace_tdse (:1385-1491) and canonical_ortho (:1495-1606): S is
    diagonalised and eigenvalues |sigma| <= 1e-6 are dropped (X = U sigma^-1/2).
    H' = X^T H X is diagonalised, and c' = exp(-H' dtau) X^+ e_first with
    X^+ = sigma^1/2 U^T is normalised.  psi = sum_i (X c')_i l_i over the
    non-extra vectors.
  * Then psi is normalised.  The reference's loop before that subtracts
    (psi.vec_conv(:,s)) vec_conv(:,s); vec_conv(:,s) is zero until state s
    converges, so the loop changes nothing.
  * residual_1vec (:1610-1651): E = psi.H psi, residual = |H psi - E psi|.
    Convergence when residual <= eps.

Pinned in tests/test_oracle_pins.py by the following checks.
  * Each of the first timesteps propagates psi to exp(-H dtau) psi /
    |exp(-H dtau) psi| (scipy.linalg.expm), to the Krylov truncation error.
  * Steps 2-4 of that check go through the inter-timestep matrix elements
    and a subspace collapse, so the same check covers them.
  * The relaxed energies and vectors are the lowest eigenpairs of a dense
    eigendecomposition.
"""
from __future__ import annotations

import numpy as np


def _canonical(S, Hs, first, tol=1e-6):
    sig, U = np.linalg.eigh(S)
    keep = np.abs(sig) > tol
    nnull = int((~keep).sum())
    # the reference counts |sigma| > tol and takes the last nlin columns
    nlin = int(keep.sum())
    U = U[:, nnull:]
    s = sig[nnull:]
    X = U / np.sqrt(s)
    Xinv = (U * np.sqrt(s)).T
    H1 = X.T @ Hs @ X
    return X, H1, Xinv[:, first], nlin


def relax(matvec, guess: np.ndarray, krydim=10, maxdim=-1, maxiter=100, step=10.0, eps=1e-7,
          trace=False):
    """``matvec(x)`` returns H x; guess: (n, nstates) initial vectors.

    Returns energies, vectors (n, nstates), residuals, iteration counts and,
    with ``trace``, per state the list of (energy, residual) of every step and
    the propagated vectors."""
    guess = np.asarray(guess, dtype=np.float64)
    n, nstates = guess.shape
    if maxdim < 0:
        maxdim = 2 * krydim
    if maxdim % krydim != 0:
        raise ValueError("maxvec must be an integer multiple of krydim")
    maxvec = maxdim + maxdim // krydim
    nsteps_cap = maxvec // (krydim + 1)
    ener = np.zeros(nstates)
    vecs = np.zeros((n, nstates))
    resid = np.zeros(nstates)
    iters = np.zeros(nstates, dtype=int)
    conv = []
    traces = []
    for s in range(nstates):
        lanc = np.zeros((n, maxvec))
        subh = np.zeros((maxvec, maxvec))
        subs = np.zeros((maxvec, maxvec))
        alph = np.zeros((krydim, nsteps_cap))
        beth = np.zeros((krydim, nsteps_cap))
        vec0 = guess[:, s].copy()
        ncurr = 0
        hpsi = matvec(vec0)
        e = vec0 @ hpsi
        res = np.linalg.norm(hpsi - e * vec0)
        tr = {"steps": [(e, res)], "psi": []}
        converged = False
        for it in range(1, maxiter + 1):
            if ncurr * (krydim + 1) == maxvec:
                kk = krydim + 1
                lanc[:, : maxvec - kk] = lanc[:, kk:].copy()
                subs[: maxvec - krydim, : maxvec - krydim] = subs[krydim:, krydim:].copy()
                subh[: maxvec - krydim, : maxvec - krydim] = subh[krydim:, krydim:].copy()
                alph[:, : nsteps_cap - 1] = alph[:, 1:].copy()
                beth[:, : nsteps_cap - 1] = beth[:, 1:].copy()
                ncurr -= 1
            ncurr += 1
            # get_lancvecs_current
            k1 = (ncurr - 1) * (krydim + 1)
            l0 = vec0.copy()
            for c in conv:
                l0 = l0 - (c @ l0) * c
            l0 = l0 / np.sqrt(l0 @ l0)
            lanc[:, k1] = l0
            q = l0
            r = matvec(q)
            alpha = np.zeros(krydim)
            beta = np.zeros(krydim)
            alpha[0] = q @ r
            r = r - alpha[0] * q
            beta[0] = np.sqrt(r @ r)
            for j1 in range(1, krydim):
                v = q
                q = r / beta[j1 - 1]
                lanc[:, k1 + j1] = q
                r = matvec(q) - beta[j1 - 1] * v
                alpha[j1] = q @ r
                r = r - alpha[j1] * q
                beta[j1] = np.sqrt(r @ r)
            lanc[:, k1 + krydim] = r / beta[krydim - 1]
            b0 = (ncurr - 1) * krydim
            for i in range(krydim):
                subh[b0 + i, b0 + i] = alpha[i]
                subs[b0 + i, b0 + i] = 1.0
            for i in range(krydim - 1):
                subh[b0 + i, b0 + i + 1] = beta[i]
                subh[b0 + i + 1, b0 + i] = beta[i]
            alph[:, ncurr - 1] = alpha
            beth[:, ncurr - 1] = beta
            # subspace_matrices
            for nprev in range(ncurr - 1):
                for ii in range(krydim):
                    i = nprev * (krydim + 1) + ii
                    i1 = nprev * krydim + ii
                    for k in range(krydim):
                        j = k1 + k
                        j1 = b0 + k
                        sj = lanc[:, i] @ lanc[:, j]
                        sj1 = lanc[:, i] @ lanc[:, j + 1]
                        if k == 0:
                            hij = alpha[k] * sj + beta[k] * sj1
                        else:
                            hij = beta[k - 1] * (lanc[:, i] @ lanc[:, j - 1]) + alpha[k] * sj + beta[k] * sj1
                        subh[i1, j1] = subh[j1, i1] = hij
                        subs[i1, j1] = subs[j1, i1] = sj
            # solve_subspace_tdse
            nvec = ncurr * krydim
            X, H1, c0, _ = _canonical(subs[:nvec, :nvec], subh[:nvec, :nvec], b0)
            lam, V = np.linalg.eigh(H1)
            c1 = (V * np.exp(-lam * step)) @ (V.T @ c0)
            c1 /= np.sqrt(c1 @ c1)
            coeff = X @ c1
            cols = [bk * (krydim + 1) + i for bk in range(ncurr) for i in range(krydim)]
            psi = lanc[:, cols] @ coeff
            psi = psi / np.sqrt(psi @ psi)
            hpsi = matvec(psi)
            e = psi @ hpsi
            res = np.sqrt(((hpsi - e * psi) ** 2).sum())
            vec0 = psi
            tr["steps"].append((e, res))
            if trace:
                tr["psi"].append(psi.copy())
            if res <= eps:
                converged = True
                iters[s] = it
                break
        if not converged:
            iters[s] = maxiter
        ener[s], vecs[:, s], resid[s] = e, vec0, res
        conv.append(vec0.copy())
        traces.append(tr)
        if not converged:
            break
    out = {"energies": ener, "vectors": vecs, "residuals": resid, "iterations": iters}
    if trace:
        out["trace"] = traces
    return out


def ondiag_guess(hii: np.ndarray, nstates: int) -> np.ndarray:
    """Unit vectors on the nstates smallest |H_jj| (initvec_ondiag,
    source/propagation/relaxation.f90:875-944; ties in index order)."""
    idx = np.argsort(np.abs(hii), kind="stable")[:nstates]
    g = np.zeros((hii.size, nstates))
    g[idx, np.arange(nstates)] = 1.0
    return g


# ---------------------------------------------------------------- Bulirsch-Stoer
class BSStepper:
    """BSStep (source/propagation/bslib.f90) with ModMidpoint, PolyExtrapol
    and AbsBSError, for a real right-hand side f(psi) = -H psi (the
    relaxation case, where the complex arithmetic keeps zero imaginary
    parts).  The Fortran ``save`` variables (OptimalColumn, NumberOfSteps,
    IntAccuracy, NewTime, Expense, Correction, OldAccuracy, OldIntOrder,
    CallIsFirstCall and PolyExtrapol's OldSquaredStepSize) live on the
    object, so they persist from call to call as in the reference."""

    MAXORDER = 16

    def __init__(self):
        M = self.MAXORDER
        self.optimal_column = 0
        self.nsteps = np.zeros(M + 2, dtype=np.int64)      # 1-based
        self.int_accuracy = 0.0
        self.new_time = 0.0
        self.expense = np.zeros(M + 2)
        self.correction = np.zeros((M + 1, M + 1))          # 1-based
        self.old_accuracy = -1.0
        self.old_int_order = -1
        self.first_call = True
        self.old_sq = np.zeros(M + 1)                        # PolyExtrapol save
        self.nfunc = 0

    def _midpoint(self, f, saved, dtpsi, large, nsteps):
        h = large / nsteps
        psi1 = saved.copy()
        psi2 = saved + h * dtpsi
        est = f(psi2)
        self.nfunc += 1
        for _ in range(2, nsteps + 1):
            swap = psi1 + 2.0 * h * est
            psi1 = psi2
            psi2 = swap
            est = f(psi2)
            self.nfunc += 1
        return 0.5 * (psi1 + psi2 + h * est)

    def _extrapol(self, column, psi, tableau, sq, int_order):
        self.old_sq[column] = sq
        delta_psi = psi.copy()
        if column == 1:
            tableau[:, 1] = psi
            return psi, delta_psi
        diff = psi.copy()
        for p in range(1, column):
            help_ = 1.0 / (self.old_sq[column - p] - sq)
            factor = sq * help_
            old_factor = self.old_sq[column - p] * help_
            interim = tableau[:, p].copy()
            tableau[:, p] = delta_psi
            delta = diff - interim
            delta_psi = factor * delta
            diff = old_factor * delta
            psi = psi + delta_psi
        if column < int_order:
            tableau[:, column] = delta_psi
        return psi, delta_psi

    def step(self, f, psi, dtpsi, int_period, abs_time, int_order, large_step, tol_error):
        """Returns (psi, int_order, actual_large_step, next_large_step, error_code)."""
        RelativeMinStep, MaxScale, ErrorSafety = 1.0e-10, 0.1, 0.25
        ReductionSafety, MinReduction, MaxReduction, Tiny = 0.7, 0.5, 1.0e-5, 1.0e-25
        if int_order < 2 or int_order > self.MAXORDER:
            return psi, int_order, 0.0, 0.0, 1
        min_large = RelativeMinStep * int_period
        next_large = -Tiny
        self.new_time = -Tiny
        if tol_error != self.old_accuracy or int_order != self.old_int_order:
            self.old_accuracy = tol_error
            self.int_accuracy = ErrorSafety * tol_error
            self.old_int_order = int_order
            for p in range(1, int_order + 2):
                self.nsteps[p] = 2 * p
            self.expense[1] = self.nsteps[1] + 1
            for p in range(2, int_order + 2):
                self.expense[p] = self.expense[p - 1] + self.nsteps[p]
            for p in range(2, int_order + 1):
                for p1 in range(1, p):
                    self.correction[p1, p] = self.int_accuracy ** (
                        (self.expense[p1 + 1] - self.expense[p + 1])
                        / ((2 * p1 + 1) * (self.expense[p + 1] - self.expense[1] + 1.0)))
            oc = 2
            while oc <= int_order - 1:                       # Do OptimalColumn = 2,IntOrder-1
                if self.expense[oc + 1] > self.expense[oc] * self.correction[oc - 1, oc]:
                    break
                oc += 1
            self.optimal_column = oc
            int_order = oc
        actual = large_step
        reduced = False
        min_work = 1.0e35
        n = psi.size
        saved = psi.copy()
        tableau = np.zeros((n, int_order + 1))               # AuxPsi(:,4:), columns 1..IntOrder-1
        if actual != next_large or abs_time != self.new_time:
            self.first_call = True
            self.optimal_column = int_order
        err = np.zeros(self.MAXORDER)
        while True:                                           # 300
            reduction = 0.0
            converged = False
            p1 = 0
            for p in range(1, int_order + 1):
                self.new_time = abs_time + actual
                sq = (actual / self.nsteps[p]) ** 2
                if actual < min_large and self.new_time < int_period:
                    return psi, int_order, actual, next_large, 2
                psi = self._midpoint(lambda x: f(x), saved, dtpsi, actual, int(self.nsteps[p]))
                psi, delta_psi = self._extrapol(p, psi, tableau, sq, int_order)
                max_error = 0.0
                p1 = 0
                if p >= 2:
                    max_error = float(np.abs(delta_psi).max())
                    p1 = p - 1
                    err[p1] = (max_error / self.int_accuracy) ** (1.0 / (2 * p1 + 1.0))
                if p >= 2 and (p >= self.optimal_column - 1 or self.first_call):
                    if max_error < tol_error:
                        converged = True
                        break
                    oc = self.optimal_column
                    if p == int_order or p == oc + 1:
                        reduction = ReductionSafety / err[p1]
                        break
                    elif p == oc:
                        if self.correction[oc - 1, oc] < err[p1]:
                            reduction = 1.0 / err[p1]
                            break
                    elif oc == int_order:
                        if self.correction[p1, int_order - 1] < err[p1]:
                            reduction = ReductionSafety * self.correction[p1, int_order - 1] / err[p1]
                            break
                    elif self.correction[p1, oc] < err[p1]:
                        reduction = self.correction[p1, oc - 1] / err[p1]
                        break
            if converged:
                break
            reduction = max(min(reduction, MinReduction), MaxReduction)   # 400
            actual = actual * reduction
            reduced = True
        self.first_call = False                               # 500
        scale = 0.0
        for p2 in range(1, p1 + 1):
            wf = max(err[p2], MaxScale)
            work = wf * self.expense[p2 + 1]
            if work < min_work:
                scale = wf
                min_work = work
                self.optimal_column = p2 + 1
        next_large = actual / scale
        oc = self.optimal_column
        if oc >= p and oc < int_order and not reduced:
            wf = max(scale / self.correction[oc - 1, oc], MaxScale)
            if self.expense[oc + 1] * wf <= min_work:
                next_large = actual / wf
                self.optimal_column = oc + 1
        return psi, int_order, actual, next_large, 0


def bs_relax(matvec, guess: np.ndarray, maxiter=100, step=10.0, eps=1e-7, toler=1e-5, log=None):
    """bs_rlx (source/propagation/relaxation.f90:315-495): each state from its
    initial vector, timesteps of length ``step`` integrated by BSStep (maximum
    order 16) until the whole step is covered; then the no-op orthogonalisation
    loop (vec_conv(:,s)), normalisation and the residual test.  tfinal =
    maxiter * step and int(tfinal / step) timesteps (evaluated in floating
    point, as the reference does).  Returns dict(eigval, eigvec, niter,
    residual, nfunc)."""
    guess = np.asarray(guess, dtype=np.float64)
    n, ns = guess.shape
    f = lambda x: -matvec(x)
    bs = BSStepper()
    intorder = 16
    tfinal = maxiter * step
    nt = int(tfinal / step)
    tiny = 1e-9
    time = 0.0
    E = np.zeros(ns)
    V = np.zeros((n, ns))
    niter = np.zeros(ns, dtype=np.int64)
    resid = np.zeros(ns)
    nextstep = 0.0
    for s in range(ns):
        psi = guess[:, s].copy()
        it_done = nt
        ok = False
        for i in range(1, nt + 1):
            inttime = 0.0
            while True:
                if inttime == 0.0:
                    stepsize = step
                elif inttime + nextstep > step:
                    stepsize = step - inttime
                else:
                    stepsize = nextstep
                dtpsi = f(psi)
                psi, intorder, true_step, nextstep, code = bs.step(f, psi, dtpsi, step, time, intorder,
                                                                   stepsize, toler)
                if code != 0:
                    raise RuntimeError(f"Bulirsch-Stoer error {code}")
                time += true_step
                inttime += true_step
                if not abs(step - inttime) > abs(tiny * step):
                    break
            psi = psi / np.sqrt(psi @ psi)
            hpsi = matvec(psi)
            e = psi @ hpsi
            r = np.linalg.norm(hpsi - e * psi)
            if log is not None:
                log.append((s, i, e, r))
            if r <= eps:
                ok = True
                it_done = i
                break
        E[s], V[:, s], niter[s], resid[s] = e, psi, it_done, r
        if not ok:
            break
    return {"eigval": E, "eigvec": V, "niter": niter, "residual": resid, "nfunc": bs.nfunc}


# ---------------------------------------------------------------- RKF45
class RKF45:
    """r8_rkf45 with r8_fehl (source/propagation/rkf45rlxlib.f90:87-795, the
    Watts-Shampine RKF45 driver) for dy/dt = f(y) = -H y.  The Fortran
    ``save`` variables (abserr_save, h, flag_save, init, kflag, kop, nfe,
    relerr_save) are members, so they persist from call to call."""

    MAXNFE = 30000

    def __init__(self):
        self.abserr_save = -1.0
        self.h = -1.0
        self.flag_save = -1000
        self.init = -1000
        self.kflag = -1000
        self.kop = -1
        self.nfe = -1
        self.relerr_save = -1.0
        self.calls = 0

    def _f(self, f, y):
        self.calls += 1
        return f(y)

    def fehl(self, f, y, h, yp):
        ch = h / 4.0
        f1 = self._f(f, y + ch * yp)
        ch = 3.0 * h / 32.0
        f2 = self._f(f, y + ch * (yp + 3.0 * f1))
        ch = h / 2197.0
        f3 = self._f(f, y + ch * (1932.0 * yp + (7296.0 * f2 - 7200.0 * f1)))
        ch = h / 4104.0
        f4 = self._f(f, y + ch * ((8341.0 * yp - 845.0 * f3) + (29440.0 * f2 - 32832.0 * f1)))
        ch = h / 20520.0
        f1n = y + ch * ((-6080.0 * yp + (9295.0 * f3 - 5643.0 * f4)) + (41040.0 * f1 - 28352.0 * f2))
        f5 = self._f(f, f1n)
        ch = h / 7618050.0
        s = y + ch * ((902880.0 * yp + (3855735.0 * f3 - 1371249.0 * f4)) + (3953664.0 * f2 + 277020.0 * f5))
        return s, f2, f3, f4, f5

    def step(self, f, y, yp, t, tout, relerr, abserr, flag):
        """One call; returns (y, yp, t, flag).  y and yp are the caller's
        arrays (yp is only recomputed where the reference recomputes it)."""
        eps = np.finfo(np.float64).eps
        if relerr < 0.0 or abserr < 0.0 or flag == 0 or flag > 8 or flag < -2:
            return y, yp, t, 8
        mflag = abs(flag)
        if mflag != 1:
            if t == tout and self.kflag != 3:
                return y, yp, t, 8
            if mflag == 2:
                if self.kflag == 3:
                    flag = self.flag_save
                    mflag = abs(flag)
                elif self.init == 0:
                    flag = self.flag_save
                elif self.kflag == 4:
                    self.nfe = 0
                elif self.kflag == 5 and abserr == 0.0:
                    raise RuntimeError("R8_RKF45: KFLAG = 5 and ABSERR = 0")
                elif self.kflag == 6 and relerr <= self.relerr_save and abserr <= self.abserr_save:
                    raise RuntimeError("R8_RKF45: KFLAG = 6")
            else:
                if flag == 3:
                    flag = self.flag_save
                    if self.kflag == 3:
                        mflag = abs(flag)
                elif flag == 4:
                    self.nfe = 0
                    flag = self.flag_save
                    if self.kflag == 3:
                        mflag = abs(flag)
                elif flag == 5 and abserr > 0.0:
                    flag = self.flag_save
                    if self.kflag == 3:
                        mflag = abs(flag)
                else:
                    raise RuntimeError("R8_RKF45: integration cannot be continued")
        self.flag_save = flag
        self.kflag = 0
        self.relerr_save = relerr
        self.abserr_save = abserr
        relerr_min = 2.0 * eps + 1.0e-12
        if relerr < relerr_min:
            self.kflag = 3
            return y, yp, t, 3
        dt = tout - t
        if mflag == 1:
            self.init = 0
            self.kop = 0
            yp = self._f(f, y)
            self.nfe = 1
            if t == tout:
                return y, yp, t, 2
        if self.init == 0:
            self.init = 1
            h = abs(dt)
            toln = 0.0
            for k in range(y.size):                      # sequential, as written
                tol = relerr * abs(y[k]) + abserr
                if tol > 0.0:
                    toln = tol
                    ypk = abs(yp[k])
                    if tol < ypk * h ** 5:
                        h = (tol / ypk) ** 0.2
            if toln <= 0.0:
                h = 0.0
            self.h = max(h, 26.0 * eps * max(abs(t), abs(dt)))
            self.flag_save = 2 if flag >= 0 else -2
        self.h = abs(self.h) if dt >= 0 else -abs(self.h)
        if 2.0 * abs(dt) <= abs(self.h):
            self.kop += 1
        if self.kop == 10000:
            self.kop = 0
            return y, yp, t, 7
        if abs(dt) <= 26.0 * eps * abs(t):
            y = y + dt * yp
            yp = self._f(f, y)
            self.nfe += 1
            return y, yp, tout, 2
        output = False
        scale = 2.0 / relerr
        ae = scale * abserr
        while True:
            hfaild = False
            hmin = 26.0 * eps * abs(t)
            dt = tout - t
            if not 2.0 * abs(self.h) <= abs(dt):
                if abs(dt) <= abs(self.h):
                    output = True
                    self.h = dt
                else:
                    self.h = 0.5 * dt
            while True:
                if self.nfe > self.MAXNFE:
                    self.kflag = 4
                    return y, yp, t, 4
                s, f2, f3, f4, f5 = self.fehl(f, y, self.h, yp)
                self.nfe += 5
                et = np.abs(y) + np.abs(s) + ae
                if np.any(et <= 0.0):
                    return y, yp, t, 5
                ee = np.abs((-2090.0 * yp + (21970.0 * f3 - 15048.0 * f4)) + (22528.0 * f2 - 27360.0 * f5))
                eeoet = float(np.max(ee / et))
                esttol = abs(self.h) * eeoet * scale / 752400.0
                if esttol <= 1.0:
                    break
                hfaild = True
                output = False
                sc = 0.9 / esttol ** 0.2 if esttol < 59049.0 else 0.1
                self.h = sc * self.h
                if abs(self.h) < hmin:
                    self.kflag = 6
                    return y, yp, t, 6
            t = t + self.h
            y = s
            yp = self._f(f, y)
            self.nfe += 1
            sc = 0.9 / esttol ** 0.2 if 0.0001889568 < esttol else 5.0
            if hfaild:
                sc = min(sc, 1.0)
            hh = max(sc * abs(self.h), hmin)
            self.h = hh if self.h >= 0 else -hh
            if output:
                return y, yp, tout, 2
            if flag <= 0:
                return y, yp, t, -2


def rkf45_relax(matvec, guess: np.ndarray, maxiter=100, step=10.0, eps=1e-7, toler=1e-5, log=None):
    """rkf45_rlx (source/propagation/relaxation.f90:501-621): each state from
    its initial vector, r8_rkf45 called with tout = i*step (relerr = abserr =
    toler), flag 1 at the start of each state; then the no-op orthogonalisation
    loop, normalisation of y (the derivative yp is not recomputed, as in the
    reference) and the residual test."""
    guess = np.asarray(guess, dtype=np.float64)
    n, ns = guess.shape
    f = lambda x: -matvec(x)
    rk = RKF45()
    E = np.zeros(ns)
    V = np.zeros((n, ns))
    niter = np.zeros(ns, dtype=np.int64)
    resid = np.zeros(ns)
    yp = np.zeros(n)
    for s in range(ns):
        psi = guess[:, s].copy()
        t = 0.0
        flag = 1
        ok = False
        it_done = maxiter
        for i in range(1, maxiter + 1):
            psi, yp, t, flag = rk.step(f, psi, yp, t, i * step, toler, toler, flag)
            if flag != 2:
                raise RuntimeError(f"Error in the RKF45 integrator. Error code: {flag}")
            psi = psi / np.sqrt(psi @ psi)
            hpsi = matvec(psi)
            e = psi @ hpsi
            r = np.linalg.norm(hpsi - e * psi)
            if log is not None:
                log.append((s, i, e, r))
            if r <= eps:
                ok = True
                it_done = i
                break
        E[s], V[:, s], niter[s], resid[s] = e, psi, it_done, r
        if not ok:
            break
    return {"eigval": E, "eigvec": V, "niter": niter, "residual": resid, "nfunc": rk.calls}
