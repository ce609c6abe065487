"""ADC(2) ground-to-excited transition density and natural transition
orbitals.  Test oracle only.

* ``trdens_vo`` restates adc2_trdens_gs_virt_occ
  (source/adclib/density_matrix.f90:497-636) for real vectors, closed shell,
  C1 symmetry, with (pq|rs) = sum_P B[P,p,q] B[P,r,s] and X[a, i] the 1h1p
  amplitudes (configuration a*o + i of oracle/csf.py; the 2h2p amplitudes do
  not enter).  Its five contributions to trdens(a, i), term by term as the
  loops sum them:
    zeroth order (:525-535)   -sqrt2 X[a,i]
    first order (:540-556)    +sqrt2 sum_{bj} (2 (bj|ai) - (bi|aj)) / (e_a + e_b - e_i - e_j) X[b,j]
    second order 1 (:561-572) +sqrt2/4 sum_b rho(a,b) X[b,i]
    second order 2 (:577-589) -sqrt2/4 sum_j rho(i,j) X[a,j]
    second order 3 (:594-627) -sqrt2/2 sum_{cj} (2 (ia|jc) - (ic|ja)) / (e_i + e_j - e_a - e_c) lambda[c,j],
                              lambda[c,j] = sum_{kb} (2 (bk|cj) - (bj|ck)) / (e_j + e_k - e_b - e_c) X[b,k]
  with rho the MP2 density of rho_mp2 as oracle/mp2.py assembles it (the
  reference's tdadc2_nto_init and adc2_nto_gs pass rho_mp2's output).
* ``tdnto`` restates tdadc2_nto (source/adclib/nto.f90:382-547): the IS part
  of the wavefunction normalised, the transition densities of its real and
  imaginary parts combined, the SVD of the nocc x nvirt transpose, and the
  pairs with sigma^2 >= 0.001 kept.

Pinned in tests/test_oracle_pins.py: with rho replaced by the F-vector's own
MP2 intermediates (rho_vv(a,c) = -4 tauA(c,a), rho_oo(k,l) = 4 tauB(l,k)),
trdens_vo(x)[a, i] equals x . f(E_ai) for every (a, i), f the 1h1p block of
the F-vector oracle (oracle/fvec.py, pinned against the exact intermediate-
state representation) of the operator whose only element is dpl(a, i) = 1:
its terms with dpl_vo (dpl(a,k), F0, FA, FB, FC) are the five above, and
every other term needs dpl_oo, dpl_vv or dpl_ov (test_oracle_trdens_is_fvector_
transpose: a dropped term, a wrong factor or a transposed operand fails it);
rho is mp2.density, pinned with the MP2 oracle.  The NTOs reconstruct the
density (U diag(sigma) V^H = trdens^T) and their count follows the threshold
(test_oracle_tdnto).
"""
from __future__ import annotations

import numpy as np

from oracle import mp2


def trdens_vo(sys, x, rho=None, V=None):
    """trdens(a, i) (nvirt x nocc) of a real vector x (ndim, 1h1p first)."""
    o, v = sys.nocc, sys.nvirt
    e = np.asarray(sys.eps, dtype=np.float64)
    eo, ev = e[:o], e[o:]
    if V is None:
        V = mp2.integrals(sys)
    if rho is None:
        rho = mp2.density(sys, V)[0]
    O, W = np.arange(o), o + np.arange(v)
    X = np.asarray(x, dtype=np.float64)[: o * v].reshape(v, o)      # X[a, i] = x[a*o + i]
    VWOWO = V[np.ix_(W, O, W, O)]                                    # (a i | b j)
    s2 = np.sqrt(2.0)
    T = -s2 * X
    # first order: (2 (bj|ai) - (bi|aj)) / (e_a + e_b - e_i - e_j)
    Dabij = ev[:, None, None, None] + ev[None, :, None, None] - eo[None, None, :, None] - eo[None, None, None, :]
    K = 2.0 * np.einsum("bjai->abij", VWOWO) - np.einsum("biaj->abij", VWOWO)
    T = T + s2 * np.einsum("abij,abij,bj->ai", K, 1.0 / Dabij, X)
    # second order 1 and 2
    T = T + 0.25 * s2 * rho[np.ix_(W, W)] @ X
    T = T - 0.25 * s2 * np.einsum("ij,aj->ai", rho[np.ix_(O, O)], X)
    # second order 3: lambda[c,j], then the contraction with (2 (ia|jc) - (ic|ja))
    Djkbc = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    L = 2.0 * np.einsum("bkcj->jkbc", VWOWO) - np.einsum("bjck->jkbc", VWOWO)
    lam = np.einsum("jkbc,jkbc,bk->cj", L, 1.0 / Djkbc, X)
    VOWOW = V[np.ix_(O, W, O, W)]                                    # (i a | j c)
    Dijac = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    Q = 2.0 * np.einsum("iajc->ijac", VOWOW) - np.einsum("icja->ijac", VOWOW)
    T = T - 0.5 * s2 * np.einsum("ijac,ijac,cj->ai", Q, 1.0 / Dijac, lam)
    return T


def tdnto(sys, wf, rho=None, thresh=1e-3):
    """tdadc2_nto for the IS part wf (ndim complex): returns (trdens (nvirt x
    nocc complex), sigma (nocc), U (nocc x nocc), V (nvirt x nvirt), npair)."""
    wf = np.asarray(wf, dtype=np.complex128)
    wf1 = wf / np.sqrt(np.vdot(wf, wf).real)
    T = trdens_vo(sys, wf1.real, rho) + 1j * trdens_vo(sys, wf1.imag, rho)
    U, sigma, VT = np.linalg.svd(T.T, full_matrices=True)
    npair = int(np.count_nonzero(sigma ** 2 >= thresh))
    return T, sigma, U, VT.conj().T, npair
