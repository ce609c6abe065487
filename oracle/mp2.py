"""MP2 ground-state quantities.  Test oracle only.

Restates source/adclib/mp2.f90 for a closed shell without frozen core
(occ1 = 1) and C1 symmetry, with (pq|rs) = sum_P B[P,p,q] B[P,r,s]:

* mp2_ener (:68-154):
  E = sum_{abij} (ia|jb) [2 (ia|jb) - (ib|ja)] / (e_i + e_j - e_a - e_b);
* mp2_d2 (:161-276) with t2amp (:280-295),
  t(i,j,a,b) = (ia|jb) / (e_i + e_j - e_a - e_b):
  d2 = max( sqrt(lmax(T^o T^o)), sqrt(lmax(T^v T^v)) ) with
  (T^o T^o)_ij = sum_{kab} t(i,k,a,b) t(j,k,a,b) and
  (T^v T^v)_ab = sum_{ijc} t(i,j,a,c) t(i,j,b,c);
* rho_mp2 (:301-413) with rhogs2_uu (:421-469), rhogs2_oo (:477-523) and
  rhogs2_ou (:531-596), assembled as the reference's loops leave it.  The oo
  block is 2 delta_ij + oo(i,j) on the diagonal, but oo(i,j) + oo(j,i) off it:
  the loop adds oo(i,j) and mirrors, then adds oo(j,i) to the mirrored value.
  The uu and ou blocks are assigned once and mirrored.

Pinned in tests/test_oracle_pins.py by independent routes:
* the energy against the spin-orbital formula
  1/4 sum |<ij||ab>|^2 / (e_i + e_j - e_a - e_b);
* d2 against the largest singular values of the amplitude matricisations;
* the second-order density correction is traceless (oo trace = -uu trace),
  and the oo/uu blocks equal the closed-shell textbook forms
  P_ij = -2 sum t_ik^ab (2 t_jk^ab - t_jk^ba) and
  P_ab = 2 sum t_ij^ac (2 t_ij^bc - t_ij^cb).
* every block, the ou block included, and the energy against exact
  Rayleigh-Schroedinger perturbation theory in the determinant space of the
  same Hamiltonian (tests/detspace.py, test_oracle_mp2_density_brute_force).
"""
from __future__ import annotations

import numpy as np


def integrals(sys):
    """(pq|rs) as a full nbas^4 array (small systems only)."""
    B = np.asarray(sys.B, dtype=np.float64)
    return np.einsum("Ppq,Prs->pqrs", B, B, optimize=True)


def _split(sys):
    o, v = sys.nocc, sys.nvirt
    e = np.asarray(sys.eps, dtype=np.float64)
    return o, v, e[:o], e[o:]


def energy(sys, V=None):
    o, v, eo, ev = _split(sys)
    V = integrals(sys) if V is None else V
    ovov = V[:o, o:, :o, o:]                       # (ia|jb)[i,a,j,b]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return float((ovov * (2.0 * ovov - ovov.transpose(0, 3, 2, 1)) / D).sum())


def amplitudes(sys, V=None):
    """t[i,j,a,b] = (ia|jb) / (e_i + e_j - e_a - e_b)."""
    o, v, eo, ev = _split(sys)
    V = integrals(sys) if V is None else V
    ovov = V[:o, o:, :o, o:]
    D = eo[:, None, None, None] - ev[None, :, None, None] + eo[None, None, :, None] - ev[None, None, None, :]
    return (ovov / D).transpose(0, 2, 1, 3)


def d2(sys, V=None):
    t = amplitudes(sys, V)
    toto = np.einsum("ikab,jkab->ij", t, t)
    tvtv = np.einsum("ijac,ijbc->ab", t, t)
    return float(max(np.sqrt(np.linalg.eigvalsh(toto)[-1]), np.sqrt(np.linalg.eigvalsh(tvtv)[-1])))


def density_frozen(sys, frozen):
    """rho_mp2 with lifrzcore (mp2.f90:333-337 and the occ1 loops of
    rhogs2_uu, rhogs2_oo and rhogs2_ou, :435-449, :491-502, :545-576):
    every occupied sum runs over the active orbitals, the frozen rows keep
    2 on the diagonal and nothing else.  That is the density of the system
    without the frozen orbitals (same integrals, same orbital energies),
    embedded; frozen holds 0-based occupied indices."""
    from types import SimpleNamespace
    o, v = sys.nocc, sys.nvirt
    frz = sorted(set(int(f) for f in frozen))
    keep = [i for i in range(o) if i not in frz] + list(range(o, o + v))
    B = np.asarray(sys.B, dtype=np.float64)
    red = SimpleNamespace(nocc=o - len(frz), nvirt=v, naux=B.shape[0], eps=np.asarray(sys.eps)[keep],
                          B=B[:, keep][:, :, keep])
    rho = np.zeros((o + v, o + v))
    for f in frz:
        rho[f, f] = 2.0
    rho[np.ix_(keep, keep)] = density(red)[0]
    return rho


def density(sys, V=None):
    """rho (nbas x nbas) of rho_mp2, with the reference's assembly."""
    o, v, eo, ev = _split(sys)
    V = integrals(sys) if V is None else V
    nb = o + v
    O, W = slice(0, o), slice(o, nb)
    vovo, ovov = V[W, O, W, O], V[O, W, O, W]
    ovvv, ooov, ovoo = V[O, W, W, W], V[O, O, O, W], V[O, W, O, O]
    ein = lambda s, *ops: np.einsum(s, *ops, optimize=True)
    # rhogs2_uu(a,b) = sum_ijc d_ijac d_ijbc [(ai|cj)(2(ib|jc) - (ic|jb)) + (ci|aj)(2(ic|jb) - (ib|jc))],
    # d_ijac = 1 / (e_a + e_c - e_i - e_j) = dv[a,i,c,j]
    dv = 1.0 / (ev[:, None, None, None] + ev[None, None, :, None] - eo[None, :, None, None] - eo[None, None, None, :])
    uu = (2.0 * ein("aicj,bicj,aicj,ibjc->ab", dv, dv, vovo, ovov)
          - ein("aicj,bicj,aicj,icjb->ab", dv, dv, vovo, ovov)
          + 2.0 * ein("aicj,bicj,ciaj,icjb->ab", dv, dv, vovo, ovov)
          - ein("aicj,bicj,ciaj,ibjc->ab", dv, dv, vovo, ovov))
    # rhogs2_oo(i,j) = -sum_kab d_ikab d_jkab [(ai|bk)(2(ja|kb) - (jb|ka)) + (bi|ak)(2(jb|ka) - (ja|kb))],
    # d_ikab = 1 / (e_a + e_b - e_i - e_k) = dv[a,i,b,k]
    oo = -(2.0 * ein("aibk,ajbk,aibk,jakb->ij", dv, dv, vovo, ovov)
           - ein("aibk,ajbk,aibk,jbka->ij", dv, dv, vovo, ovov)
           + 2.0 * ein("aibk,ajbk,biak,jbka->ij", dv, dv, vovo, ovov)
           - ein("aibk,ajbk,biak,jakb->ij", dv, dv, vovo, ovov))
    # rhogs2_ou(i,a) = -(term1 + term2) / (e_a - e_i)
    # term1 = sum_jbc d_ijbc [(ib|jc)((jb|ac) - 2(jc|ab)) + (ic|jb)((jc|ab) - 2(jb|ac))],
    # d_ijbc = 1 / (e_b + e_c - e_i - e_j) = dv[b,i,c,j]
    t1 = (ein("bicj,ibjc,jbac->ia", dv, ovov, ovvv)
          - 2.0 * ein("bicj,ibjc,jcab->ia", dv, ovov, ovvv)
          + ein("bicj,icjb,jcab->ia", dv, ovov, ovvv)
          - 2.0 * ein("bicj,icjb,jbac->ia", dv, ovov, ovvv))
    # term2 = sum_jkb d_jkab [(ji|kb)(2(ja|kb) - (jb|ka)) + (jb|ki)(2(jb|ka) - (ja|kb))],
    # d_jkab = 1 / (e_a + e_b - e_j - e_k) = dv[a,j,b,k]
    t2 = (2.0 * ein("ajbk,jikb,jakb->ia", dv, ooov, ovov)
          - ein("ajbk,jikb,jbka->ia", dv, ooov, ovov)
          + 2.0 * ein("ajbk,jbki,jbka->ia", dv, ovoo, ovov)
          - ein("ajbk,jbki,jakb->ia", dv, ovoo, ovov))
    ou = -(t1 + t2) / (ev[None, :] - eo[:, None])
    rho = np.zeros((nb, nb))
    rho[np.arange(o), np.arange(o)] = 2.0
    for i in range(o):
        for j in range(o):
            rho[i, j] = rho[i, j] + oo[i, j]
            rho[j, i] = rho[i, j]
    rho[o:, o:] = np.triu(uu) + np.triu(uu, 1).T
    rho[:o, o:] = ou
    rho[o:, :o] = ou.T
    return rho, {"oo": oo, "uu": uu, "ou": ou}
