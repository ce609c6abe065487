"""Closed-shell RHF (DIIS) and the MO quantities the ADC(2) library takes.

The reference redoes the RHF on its own integrals before the ADC step
(source/qclib/load_electronic_structure.f90:153-342); so does this module,
from the core-Hamiltonian guess, to an energy change below 1e-12 Eh and a
commutator [F, D S] below 1e-10.
"""
from __future__ import annotations

import dataclasses

import numpy as np
import scipy.linalg as sl

from ..synth import System


@dataclasses.dataclass
class RHF:
    energy: float
    eps: np.ndarray      # orbital energies, ascending
    C: np.ndarray        # AO x MO
    nocc: int
    e_nuc: float
    iterations: int


def rhf(mol, ints, maxiter: int = 200, conv: float = 1e-12) -> RHF:
    S, H = ints["S"], ints["T"] + ints["V"]
    G = ints["eri"]
    nocc = mol.nelec // 2
    e_nuc = mol.nuclear_repulsion()
    X = sl.fractional_matrix_power(S, -0.5).real
    eps, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    focks, errs = [], []
    e_old = 0.0
    for it in range(1, maxiter + 1):
        Co = C[:, :nocc]
        P = 2.0 * Co @ Co.T
        J = np.einsum("pqrs,rs->pq", G, P, optimize=True)
        K = np.einsum("prqs,rs->pq", G, P, optimize=True)
        F = H + J - 0.5 * K
        e = 0.5 * np.sum(P * (H + F)) + e_nuc
        err = F @ P @ S - S @ P @ F
        focks.append(F)
        errs.append(err)
        focks, errs = focks[-8:], errs[-8:]
        if abs(e - e_old) < conv and np.abs(err).max() < 1e-10:
            eps, Cp = np.linalg.eigh(X.T @ F @ X)
            return RHF(float(e), eps, X @ Cp, nocc, e_nuc, it)
        e_old = e
        m = len(focks)
        Bm = -np.ones((m + 1, m + 1))
        Bm[m, m] = 0.0
        for i in range(m):
            for j in range(m):
                Bm[i, j] = np.sum(errs[i] * errs[j])
        rhs = np.zeros(m + 1)
        rhs[m] = -1.0
        c = np.linalg.lstsq(Bm, rhs, rcond=None)[0]
        Fd = sum(c[i] * focks[i] for i in range(m))
        eps, Cp = np.linalg.eigh(X.T @ Fd @ X)
        C = X @ Cp
    raise RuntimeError("RHF did not converge")


def mo_integrals(ints, C):
    G = ints["eri"]
    return np.einsum("pqrs,pi,qj,rk,sl->ijkl", G, C, C, C, C, optimize=True)


def cholesky_factors(eri_mo: np.ndarray, tol: float = 1e-12) -> np.ndarray:
    """B[P,p,q] with sum_P B[P,p,q] B[P,r,s] = (pq|rs) to tol, by pivoted
    Cholesky of the matrix M[(pq),(rs)] over the pairs p >= q."""
    n = eri_mo.shape[0]
    iu, ju = np.tril_indices(n)
    M = eri_mo[iu, ju][:, iu, ju]
    d = np.diag(M).copy()
    Ls = []
    while True:
        k = int(np.argmax(d))
        if d[k] <= tol:
            break
        col = M[:, k].copy()
        for Lv in Ls:
            col -= Lv * Lv[k]
        Lv = col / np.sqrt(d[k])
        Ls.append(Lv)
        d -= Lv * Lv
        d[k] = 0.0
    L = np.array(Ls)  # (naux, npair)
    B = np.zeros((L.shape[0], n, n))
    B[:, iu, ju] = L
    B[:, ju, iu] = L
    return B


def mo_system(scf: RHF, ints, frozen: int = 0, nvirt: int | None = None, tol: float = 1e-11):
    """(System, B residual): orbitals frozen..nmo (or the first nvirt virtuals)
    as the ADC(2) system: eps and the Cholesky factors of their MO integrals."""
    nmo = scf.C.shape[1]
    hi = nmo if nvirt is None else scf.nocc + nvirt
    C = scf.C[:, frozen:hi]
    eri = mo_integrals(ints, C)
    B = cholesky_factors(eri, tol)
    res = float(np.abs(np.einsum("Ppq,Prs->pqrs", B, B, optimize=True) - eri).max())
    sys_ = System(scf.nocc - frozen, hi - scf.nocc, B.shape[0], scf.eps[frozen:hi].copy(), np.ascontiguousarray(B))
    return sys_, res
