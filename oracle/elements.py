"""ADC(2)-s matrix elements and the stored-matrix build.  Test oracle only.

Restates, term by term, the spin-free expressions of
source/adclib/adc_ph.f90 with the integral lookup
``vpqrs(p,q,r,s) = (pq|rs)`` (Mulliken, source/qclib/vpqrsmod.f90:13-27) taken
from a full 4-index array built from the density-fitting factors.  MO indices
are 0-based with the occupied orbitals first, so the 1-based reference index
``p`` is ``p-1`` here and a virtual ``a`` of the configuration tables is the
MO ``nocc + a``.

Functions follow:

* ``k_ph_ph``/``c1_ph_ph``          adc_ph.f90:31-53
* ``ca_ph_ph``                      adc_ph.f90:769-812 (all a, a')
* ``cb_ph_ph``                      adc_ph.f90:819-912 (all k, k')
* ``cc_ph_ph``                      adc_ph.f90:918-957
* ``c1..c5_ph_2p2h``                adc_ph.f90:974-1073
* ``k_2p2h``                        adc_ph.f90:1090-1096
* ``build_blocks``                  get_matrix.f90:267-311 (diagonal) and
                                    get_matrix.f90:705-1186 (off-diagonal:
                                    ca/cb precompute :804-822, ph-ph :848-884,
                                    the five coupling blocks :895-1054, the
                                    |h| > minc = 1e-12 filter,
                                    source/include/parameters.f90:134)

The result is kept as blocks (H11 dense n1 x n1, C dense n1 x n2, D the 2h2p
diagonal) rather than the reference's COO list; ``coo`` emits the upper
triangle as (value, i, j) 1-based triples in the reference's record order for
the CPU baseline.

Pinned by: the closed forms of ``test_oracle_pins`` (zero-coupling limit,
ADC(1)/CIS eigenvalues by dense eigh, symmetry), agreement with the
independent DF formulation in ``dfsigma`` and with the Lowdin-partitioned
exact eigenvalues of ``schur``; and against the exact intermediate-state
representation in the determinant space (tests/detspace.py isr_orders,
test_oracle_isr_brute_force): H11 order by order through second order
element by element, C and D through the 2h2p-basis invariants C C^T and
C D C^T.
"""
from __future__ import annotations

import numpy as np

from . import csf

MINC = 1e-12  # source/include/parameters.f90:134


def eri(B: np.ndarray) -> np.ndarray:
    """(pq|rs) = sum_P B[P,p,q] B[P,r,s]; small systems only."""
    nb = B.shape[1]
    Bm = B.reshape(B.shape[0], nb * nb)
    return (Bm.T @ Bm).reshape(nb, nb, nb, nb)


def ca_ph_ph(V, e, o, v):
    """CA(a,a') for all virtual pairs, adc_ph.f90:769-812."""
    occ = np.arange(o)
    vir = np.arange(o, o + v)
    eu = e[occ][:, None, None]
    ev = e[occ][None, :, None]
    er = e[vir][None, None, :]
    eijc = eu + ev - er                     # [u, v, r]
    ea = e[vir]
    out = np.empty((v, v))
    for ia, A in enumerate(vir):
        for ib, A1 in enumerate(vir):
            if ib < ia:
                out[ia, ib] = out[ib, ia]
                continue
            DA = (eijc - 0.5 * (ea[ia] + ea[ib])) / ((eijc - ea[ia]) * (eijc - ea[ib]))
            # term = V(a,u,r,v)*(2V(u,a1,v,r)-V(u,r,v,a1)) + V(a,v,r,u)*(2V(u,r,v,a1)-V(u,a1,v,r))
            Va_urv = V[A][np.ix_(occ, vir, occ)].transpose(0, 2, 1)      # V(a,u,r,v) -> [u,v,r]
            Vu_a1_vr = V[np.ix_(occ, [A1], occ, vir)][:, 0]              # V(u,a1,v,r) -> [u,v,r]
            Vu_r_va1 = V[np.ix_(occ, vir, occ, [A1])][..., 0].transpose(0, 2, 1)  # V(u,r,v,a1) -> [u,v,r]
            Va_vru = V[A][np.ix_(occ, vir, occ)].transpose(2, 0, 1)      # V(a,v,r,u) indexed [u,v,r]
            term = Va_urv * (2 * Vu_a1_vr - Vu_r_va1) + Va_vru * (2 * Vu_r_va1 - Vu_a1_vr)
            out[ia, ib] = -0.5 * np.sum(DA * term)
    return out


def cb_ph_ph(V, e, o, v):
    """CB(k,k') for all occupied pairs, adc_ph.f90:819-912 (the active
    "dumbed down" loop)."""
    occ = np.arange(o)
    vir = np.arange(o, o + v)
    ei = e[occ][None, None, :]
    ec = e[vir][:, None, None]
    ed = e[vir][None, :, None]
    eicd = ei - ec - ed                     # [c, d, i]
    out = np.empty((o, o))
    for k in range(o):
        for k1 in range(o):
            if k1 < k:
                out[k, k1] = out[k1, k]
                continue
            DB = (eicd + 0.5 * (e[k] + e[k1])) / ((eicd + e[k]) * (eicd + e[k1]))
            # term = V(r,k,s,u)*(2V(k1,r,u,s)-V(k1,s,u,r)) + V(r,u,s,k)*(2V(k1,s,u,r)-V(k1,r,u,s))
            V_rksu = V[np.ix_(vir, [k], vir, occ)][:, 0]            # [c,d,i]
            V_k1rus = V[k1][np.ix_(vir, occ, vir)].transpose(0, 2, 1)  # V(k1,c,i,d) -> [c,d,i]
            V_k1sur = V[k1][np.ix_(vir, occ, vir)].transpose(2, 0, 1)  # V(k1,d,i,c) -> [c,d,i]
            V_rusk = V[np.ix_(vir, occ, vir, [k])][..., 0].transpose(0, 2, 1)  # V(c,i,d,k) -> [c,d,i]
            term = V_rksu * (2 * V_k1rus - V_k1sur) + V_rusk * (2 * V_k1sur - V_k1rus)
            out[k, k1] = -0.5 * np.sum(DB * term)
    return out


def cc_ph_ph(V, e, o, v):
    """CC(ak,a'k') for all 1h1p pairs, adc_ph.f90:918-957.  Rows/cols in the
    1h1p order (a outer, k inner)."""
    occ = np.arange(o)
    vir = np.arange(o, o + v)
    # t[a,k,c,i] = 2 V(a,k,c,i) - V(a,i,c,k)
    Vakci = V[np.ix_(vir, occ, vir, occ)]
    t = 2 * Vakci - Vakci.transpose(0, 3, 2, 1)
    ek = e[occ]
    ea = e[vir]
    eic = (e[occ][None, :] - e[vir][:, None])   # [c, i]
    d = (ek[None, :] - ea[:, None])             # [a, k]
    n1 = o * v
    T = t.reshape(n1, v * o)                    # rows (a,k), cols (c,i)
    dd = d.reshape(n1)
    den = dd[:, None] + eic.reshape(1, v * o)   # e_k - e_a + e_i - e_c, [(a,k),(c,i)]
    # DC = (0.5(dk + dk') + eic) / ((dk + eic)(dk' + eic)) = 0.5(1/den + 1/den')
    inv = 1.0 / den
    return 0.5 * ((T * inv) @ T.T + T @ (T * inv).T)


def _mask(x):
    return x.astype(np.float64)


def coupling(V, o, v, typ, A, Bv, K, L, rows_a, rows_k):
    """Coupling block C[(a,k), J] for one 2h2p type; adc_ph.f90:974-1073.

    rows_a, rows_k: 1h1p virtual/occupied indices (0-based) of the rows.
    A, Bv, K, L: the config's a', b', k', l' (0-based).  Virtual indices are
    shifted to MO numbering inside.
    """
    a = (rows_a + o)[:, None]
    k = rows_k[:, None]
    ap = (A + o)[None, :]
    bp = (Bv + o)[None, :]
    kp = K[None, :]
    lp = L[None, :]
    if typ == 4:  # C1_ph_2p2h, spin case 1
        f = (-_mask(a == ap) * (V[kp, k, lp, bp] + V[kp, bp, lp, k])
             - _mask(a == bp) * (V[kp, k, lp, ap] + V[kp, ap, lp, k])
             + _mask(k == kp) * (V[a, ap, lp, bp] + V[a, bp, lp, ap])
             + _mask(k == lp) * (V[a, ap, kp, bp] + V[a, bp, kp, ap]))
        return f / np.sqrt(2.0)
    if typ == 5:  # C2_ph_2p2h, spin case 2
        f = (-_mask(a == ap) * (V[kp, k, lp, bp] - V[kp, bp, lp, k])
             + _mask(a == bp) * (V[kp, k, lp, ap] - V[kp, ap, lp, k])
             + _mask(k == kp) * (V[a, ap, lp, bp] - V[a, bp, lp, ap])
             - _mask(k == lp) * (V[a, ap, kp, bp] - V[a, bp, kp, ap]))
        return np.sqrt(3.0) * f / np.sqrt(2.0)
    if typ == 3:  # C3_ph_2p2h(j,k,ipr,kpr,lpr), a'=b'
        return (-_mask(a == ap) * (V[kp, k, lp, ap] + V[kp, ap, lp, k])
                + _mask(k == kp) * V[a, ap, lp, ap]
                + _mask(k == lp) * V[a, ap, kp, ap])
    if typ == 2:  # C4_ph_2p2h(a,k,apr,bpr,kpr), k'=l'
        return (_mask(a == ap) * V[kp, k, kp, bp]
                + _mask(a == bp) * V[kp, k, kp, ap]
                - _mask(k == kp) * (V[a, ap, kp, bp] + V[a, bp, kp, ap]))
    if typ == 1:  # C5_ph_2p2h(a,k,apr,kpr)
        return np.sqrt(2.0) * (_mask(a == ap) * V[kp, ap, kp, k]
                               - _mask(k == kp) * V[a, ap, kp, ap])
    raise ValueError(typ)


class Blocks:
    """ADC(2)-s matrix as H = [[H11, C], [C^T, diag(D)]]."""

    def __init__(self, H11, C, D, o, v, type_offsets):
        self.H11, self.C, self.D, self.o, self.v = H11, C, D, o, v
        self.type_offsets = type_offsets  # column offsets of t1..t5 in C, plus the end
        self.n1 = H11.shape[0]
        self.ndim = self.n1 + D.size

    def diag(self):
        return np.concatenate([np.diag(self.H11), self.D])

    def sigma(self, X):
        """W = H X for X of shape (ndim,) or (ndim, nvec)."""
        X1 = X[: self.n1]
        X2 = X[self.n1:]
        W1 = self.H11 @ X1 + self.C @ X2
        W2 = self.C.T @ X1 + (self.D * X2.T).T
        return np.concatenate([W1, W2], axis=0)

    def dense(self):
        n1 = self.n1
        H = np.zeros((self.ndim, self.ndim))
        H[:n1, :n1] = self.H11
        H[:n1, n1:] = self.C
        H[n1:, :n1] = self.C.T
        H[np.arange(n1, self.ndim), np.arange(n1, self.ndim)] = self.D
        return H

    def coo(self):
        """Upper-triangle off-diagonal elements in the reference record order
        (ph-ph rows i<j, then the coupling blocks type by type,
        get_matrix.f90:848-1054), as (val, i, j) with 1-based indices."""
        n1 = self.n1
        iu, ju = np.triu_indices(n1, 1)
        vals = [self.H11[iu, ju]]
        ii = [iu + 1]
        jj = [ju + 1]
        offs = self.type_offsets
        for t in range(5):
            blk = self.C[:, offs[t]:offs[t + 1]]
            # row-major nonzeros == the reference loop (i outer, j inner)
            r, c = np.nonzero(np.abs(blk) > MINC)
            vals.append(blk[r, c])
            ii.append(r + 1)
            jj.append(c + offs[t] + n1 + 1)
        val = np.concatenate(vals)
        i = np.concatenate(ii)
        j = np.concatenate(jj)
        keep = np.abs(val) > MINC
        return val[keep], i[keep].astype(np.int32), j[keep].astype(np.int32)


def build_blocks(sys, second_order: bool = True, apply_minc: bool = True) -> Blocks:
    """get_diag_adc2_save (get_matrix.f90:267-311) + get_offdiag_adc2_save_omp
    (get_matrix.f90:705-1186) restated into dense blocks."""
    o, v = sys.nocc, sys.nvirt
    e = sys.eps
    V = eri(sys.B)
    typ, A, Bv, K, L = csf.table(o, v)
    n1 = o * v
    ra, rk = A[:n1], K[:n1]
    # ph-ph block: C1 + delta_kk' ca + delta_aa' cb + CC (get_matrix.f90:848-862)
    aa = ra[:, None]
    a2 = ra[None, :]
    kk = rk[:, None]
    k2 = rk[None, :]
    H11 = 2.0 * V[aa + o, kk, a2 + o, k2] - V[aa + o, a2 + o, kk, k2]
    if second_order:
        ca = ca_ph_ph(V, e, o, v)
        cb = cb_ph_ph(V, e, o, v)
        H11 = H11 + (kk == k2) * ca[aa, a2] + (aa == a2) * cb[kk, k2]
        H11 = H11 + cc_ph_ph(V, e, o, v)
    # zeroth order on the diagonal (K_ph_ph, get_matrix.f90:289)
    H11[np.arange(n1), np.arange(n1)] += e[ra + o] - e[rk]
    if apply_minc:
        off = ~np.eye(n1, dtype=bool)
        H11[off & (np.abs(H11) <= MINC)] = 0.0
    # coupling blocks, type by type
    blocks = []
    offs = [0]
    for t in (1, 2, 3, 4, 5):
        sel = typ == t
        blocks.append(coupling(V, o, v, t, A[sel], Bv[sel], K[sel], L[sel], ra, rk))
        offs.append(offs[-1] + blocks[-1].shape[1])
    C = np.concatenate(blocks, axis=1)
    if apply_minc:
        C[np.abs(C) <= MINC] = 0.0
    # 2h2p diagonal, K_2p2h_2p2h (adc_ph.f90:1090-1096, get_matrix.f90:300)
    sel = typ > 0
    D = e[A[sel] + o] + e[Bv[sel] + o] - e[K[sel]] - e[L[sel]]
    return Blocks(H11, C, D, o, v, offs)
