"""ADC(2) F-vector f_J = <Psi_J|D|Psi_0>.  Test oracle only.

Restates get_modifiedtm_adc2 (source/adclib/get_moment.f90:153-442) for a
closed shell with C1 symmetry (every symmetry test of the term functions
passes) and the term functions of source/adclib/dipole_ph.f90, with
(pq|rs) = sum_P B[P,p,q] B[P,r,s] and dpl the nbas x nbas one-electron
operator matrix indexed as the reference indexes it (dpl[p, q] = dpl(p,q)).

1h1p block, config (a, k):
    f = -sqrt(2) * ( dpl(a,k) + F0 + F25..F210 + FA + FB + FC + F21..F24 )
with FA = sum_c tauA(c,a) dpl(c,k), FB = sum_l dpl(a,l) tauB(l,k),
F21/F22 = sum_l tau21/22(a,l) dpl(l,k), F23/F24 = sum_c dpl(a,c) tau23/24(c,k)
and FC = FC_ph_new with tauC (get_moment.f90:198-367).
2h2p block in the configuration order of oracle/csf.py:
    t1 FI_2p2h(a,k), t2 FII_2p2h(a,b,k), t3 FIII_2p2h(a,k,l),
    t4 FIV1_2p2h(a,b,k,l), t5 FIV2_2p2h(a,b,k,l) (get_moment.f90:370-435).

Two restatements:
* ``literal``: each function's sums as written, restricted sums (b < c,
  l < m) and their diagonal terms kept apart (dipole_ph.f90:24-46, 189-656,
  662-846, 896-1246);
* ``compact``: the same quantities from the amplitudes
  t[i,j,a,b] = (ia|jb) / (e_i + e_j - e_a - e_b), the restricted sums folded
  into full sums (each summand is symmetric under the exchange of the two
  restricted indices and its diagonal term is half the folded summand), and
  the 2h2p functions from one tensor
  R(a,b,k,l) = sum_m dpl(m,k) t[m,l,a,b] + sum_m dpl(m,l) t[k,m,a,b]
             - sum_c dpl(a,c) t[k,l,c,b] - sum_c dpl(b,c) t[k,l,a,c]:
  FI = -R(a,a,k,k), FII = -sqrt(2) R(a,b,k,k), FIII = sqrt(2) R(a,a,k,l),
  FIV1 = R(a,b,k,l) + R(b,a,k,l), FIV2 = sqrt(3) (R(a,b,k,l) - R(b,a,k,l)).

Pinned in tests/test_oracle_pins.py by: literal == compact term by term
(symmetric and non-symmetric dpl); the number operator (dpl = identity) gives
f = 0 exactly, since <Psi_J|N|Psi_0> = N <Psi_J|Psi_0> = 0 (this fixes the
relative signs and factors of F21..F24 and of every 2h2p function); and the
density-like intermediates against the MP2 density of oracle/mp2.py
(tauA(c,a) = -uu(a,c)/4 and tauB(m,k) = oo(k,m)/4, both pinned by the
textbook closed-shell MP2 density; tau21 + tau22 = -ou^T/2); and every term
against the exact intermediate-state representation in the determinant
space of the same Hamiltonian (tests/detspace.py isr_orders,
test_oracle_isr_brute_force): the 1h1p block order by order through second
order element by element (the reference's 1h1p configurations carry the sign
-1 against C_ak = g_ak/sqrt(2)), the 2h2p block through the invariants of the
2h2p basis (C f2, |f2|^2).
"""
from __future__ import annotations

import numpy as np

from oracle import csf, mp2


def _setup(sys, dpl):
    o, v = sys.nocc, sys.nvirt
    e = np.asarray(sys.eps, dtype=np.float64)
    V = mp2.integrals(sys)
    d = np.asarray(dpl, dtype=np.float64)
    assert d.shape == (o + v, o + v)
    return o, v, e[:o], e[o:], V, d


# ---------------------------------------------------------------- literal
def literal_terms(sys, dpl):
    """Every 1h1p term (v x o arrays [a, k]) and 2h2p function, restated."""
    o, v, eo, ev, V, d = _setup(sys, dpl)
    nb = o + v
    O, W = np.arange(o), o + np.arange(v)
    blk = lambda *ix: V[np.ix_(*ix)]
    ein = lambda s, *ops: np.einsum(s, *ops, optimize=True)
    d_vo, d_ov, d_oo, d_vv = d[np.ix_(W, O)], d[np.ix_(O, W)], d[np.ix_(O, O)], d[np.ix_(W, W)]
    lo = np.triu(np.ones((o, o)), 1)          # l < m
    lv = np.triu(np.ones((v, v)), 1)          # b < c
    VWOWO = blk(W, O, W, O)                   # (a l | b m)
    VOWOW = blk(O, W, O, W)                   # (l c | m b)
    VOWWW = blk(O, W, W, W)                   # (l b | a c)
    VWWOO = blk(W, W, O, O)                   # (a c | m l)
    VOOOW = blk(O, O, O, W)                   # (l k | m c)
    VOWOO = blk(O, W, O, O)                   # (l b | m n)
    VOOOO = blk(O, O, O, O)
    VWWWW = blk(W, W, W, W)
    t = {}

    # F0_ph (dipole_ph.f90:24-46)
    E = ev[:, None, None, None] + ev[None, :, None, None] - eo[None, None, :, None] - eo[None, None, None, :]  # [a,b,k,l]
    t["F0"] = -(ein("bl,abkl,akbl->ak", d_vo, 1.0 / E, 2.0 * VWOWO)
                - ein("bl,abkl,albk->ak", d_vo, 1.0 / E, VWOWO))

    # tauA (:896-935), FA: sum_c tauA(c,a) dpl(c,k)
    Dlmba = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]  # [l,m,b,a]
    dg = lambda x: np.einsum("llba->lba", x)  # l = m
    tauA = -0.5 * ein("lba,lbc,albl,lclb->ca", 1.0 / dg(Dlmba), 1.0 / dg(Dlmba), VWOWO, VOWOW)
    tauA += -0.5 * ein("lm,lmba,lmbc,albm,lcmb->ca", lo, 1.0 / Dlmba, 1.0 / Dlmba, VWOWO, 2.0 * VOWOW)
    tauA -= -0.5 * ein("lm,lmba,lmbc,albm,lbmc->ca", lo, 1.0 / Dlmba, 1.0 / Dlmba, VWOWO, VOWOW)
    tauA += -0.5 * ein("lm,lmba,lmbc,ambl,lbmc->ca", lo, 1.0 / Dlmba, 1.0 / Dlmba, VWOWO, 2.0 * VOWOW)
    tauA -= -0.5 * ein("lm,lmba,lmbc,ambl,lcmb->ca", lo, 1.0 / Dlmba, 1.0 / Dlmba, VWOWO, VOWOW)
    t["tauA"] = tauA
    t["FA"] = ein("ca,ck->ak", tauA, d_vo)

    # tauB (:939-981), FB: sum_l dpl(a,l) tauB(l,k)
    Dklbc = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]  # [k,l,b,c]
    dgv = lambda x: np.einsum("klbb->klb", x)  # b = c
    tauB = -0.5 * ein("klb,lmb,bkbl,mblb->mk", 1.0 / dgv(Dklbc), 1.0 / dgv(Dklbc).transpose(1, 0, 2), VWOWO, VOWOW)
    A1 = ein("bc,klbc,lmbc->klmbc", lv, 1.0 / Dklbc, 1.0 / Dklbc)  # [k,l,m,b,c]: 1/(D_klbc D_lmbc)
    tauB += -0.5 * (ein("klmbc,bkcl,mblc->mk", A1, VWOWO, 2.0 * VOWOW) - ein("klmbc,bkcl,mclb->mk", A1, VWOWO, VOWOW)
                    + ein("klmbc,blck,mclb->mk", A1, VWOWO, 2.0 * VOWOW) - ein("klmbc,blck,mblc->mk", A1, VWOWO, VOWOW))
    t["tauB"] = tauB
    t["FB"] = ein("al,lk->ak", d[np.ix_(W, O)], tauB)

    # tau21 (:985-1029): -1/(e_m - e_a) [...], F21 = sum_l tau21(a,l) dpl(l,k)
    ema = eo[None, :] - ev[:, None]           # [a,m]
    Dlmbc = Dklbc                             # same function of (l,m,b,c)
    s = (ein("lmb,blbm,lbab->am", 1.0 / dgv(Dlmbc), VWOWO, VOWWW)
         + ein("bc,lmbc,blcm,lbac->am", lv, 1.0 / Dlmbc, VWOWO, 2.0 * VOWWW)
         - ein("bc,lmbc,blcm,lcab->am", lv, 1.0 / Dlmbc, VWOWO, VOWWW)
         + ein("bc,lmbc,bmcl,lcab->am", lv, 1.0 / Dlmbc, VWOWO, 2.0 * VOWWW)
         - ein("bc,lmbc,bmcl,lbac->am", lv, 1.0 / Dlmbc, VWOWO, VOWWW))
    tau21 = -s / ema
    t["tau21"] = tau21
    t["F21"] = ein("al,lk->ak", tau21, d_oo)

    # tau22 (:1033-1076): +1/(e_n - e_a) [...]
    s = (ein("lba,lbln,blal->an", 1.0 / dg(Dlmba), VOWOO, VWOWO)
         + ein("lm,lmba,lbmn,blam->an", lo, 1.0 / Dlmba, VOWOO, 2.0 * VWOWO)
         - ein("lm,lmba,lbmn,bmal->an", lo, 1.0 / Dlmba, VOWOO, VWOWO)
         + ein("lm,lmba,lnmb,bmal->an", lo, 1.0 / Dlmba, blk(O, O, O, W), 2.0 * VWOWO)
         - ein("lm,lmba,lnmb,blam->an", lo, 1.0 / Dlmba, blk(O, O, O, W), VWOWO))
    tau22 = s / ema
    t["tau22"] = tau22
    t["F22"] = ein("al,lk->ak", tau22, d_oo)

    # tau23 (:1080-1125): -1/(e_k - e_b) [...], F23 = sum_c dpl(a,c) tau23(c,k)
    Dlmbc2 = Dlmba                            # [l,m,b,c] with e_l + e_m - e_b - e_c
    s = (ein("lbc,blcl,lklc->bk", 1.0 / dg(Dlmbc2), VWOWO, VOOOW)
         + ein("lm,lmbc,blcm,lkmc->bk", lo, 1.0 / Dlmbc2, VWOWO, 2.0 * VOOOW)
         - ein("lm,lmbc,blcm,lcmk->bk", lo, 1.0 / Dlmbc2, VWOWO, blk(O, W, O, O))
         + ein("lm,lmbc,bmcl,lcmk->bk", lo, 1.0 / Dlmbc2, VWOWO, 2.0 * blk(O, W, O, O))
         - ein("lm,lmbc,bmcl,lkmc->bk", lo, 1.0 / Dlmbc2, VWOWO, VOOOW))
    tau23 = -s / ema
    t["tau23"] = tau23
    t["F23"] = ein("ac,ck->ak", d_vv, tau23)

    # tau24 (:1129-1177): +1/(e_k - e_d) [...], F24 = sum_c dpl(a,c) tau24(c,k)
    s = (ein("klb,bkbl,dblb->dk", 1.0 / dgv(Dklbc), VWOWO, blk(W, W, O, W))
         + ein("bc,klbc,bkcl,dblc->dk", lv, 1.0 / Dklbc, VWOWO, 2.0 * blk(W, W, O, W))
         - ein("bc,klbc,bkcl,dclb->dk", lv, 1.0 / Dklbc, VWOWO, blk(W, W, O, W))
         + ein("bc,klbc,blck,dclb->dk", lv, 1.0 / Dklbc, VWOWO, 2.0 * blk(W, W, O, W))
         - ein("bc,klbc,blck,dblc->dk", lv, 1.0 / Dklbc, VWOWO, blk(W, W, O, W)))
    tau24 = s / ema
    t["tau24"] = tau24
    t["F24"] = ein("ac,ck->ak", d_vv, tau24)

    # tauC (:1181-1209) and FC_ph_new (:1213-1242)
    Dmlcb = Dklbc                             # [m,l,c,b]
    tauC = 0.5 * (ein("cm,mlcb,mclb->bl", d_vo, 1.0 / Dmlcb, 2.0 * VOWOW)
                  - ein("cm,mlcb,mblc->bl", d_vo, 1.0 / Dmlcb, VOWOW))
    t["tauC"] = tauC
    Dklab = Dklbc                             # [k,l,a,b]
    t["FC"] = (ein("bl,klab,akbl->ak", tauC, 1.0 / Dklab, 2.0 * VWOWO)
               - ein("bl,klab,albk->ak", tauC, 1.0 / Dklab, VWOWO))

    # F25 (:390-428)
    Dkmac = Dklbc                             # [k,m,a,c]
    Dlmbc = Dklbc
    A = 2.0 * blk(W, O, O, W) - blk(W, W, O, O).transpose(0, 3, 2, 1)   # (2(ak|lb) - (ab|lk)) [a,k,l,b]
    Bm = 2.0 * VWOWO - VWOWO.transpose(0, 3, 2, 1)                       # (2(bl|cm) - (bm|cl)) [b,l,c,m]
    t["F25"] = ein("mc,kmac,lmbc,aklb,blcm->ak", d_ov, 1.0 / Dkmac, 1.0 / Dlmbc, A, Bm)

    # F26 (:432-469)
    C = 2.0 * blk(O, W, W, O) - blk(O, O, W, W).transpose(0, 3, 2, 1)   # (2(lb|cm) - (lm|cb)) [l,b,c,m]
    Dk = 2.0 * VWOWO - VWOWO.transpose(0, 3, 2, 1)                       # (2(ak|bl) - (al|bk)) [a,k,b,l]
    t["F26"] = ein("mc,klab,kmac,lbcm,akbl->ak", d_ov, 1.0 / Dklab, 1.0 / Dkmac, C, Dk)

    # F27 (:473-513)
    Dkmbc = Dklbc
    P1 = ein("almc,cmbk->albkmc", blk(W, O, O, W), 2.0 * blk(W, O, W, O)) - ein("almc,ckbm->albkmc", blk(W, O, O, W), blk(W, O, W, O))
    P2 = ein("acml,ckbm->albkmc", VWWOO, 2.0 * blk(W, O, W, O)) - ein("acml,cmbk->albkmc", VWWOO, blk(W, O, W, O))
    t["F27"] = -ein("lb,klab,kmbc,albkmc->ak", d_ob(d, O, W), 1.0 / Dklab, 1.0 / Dkmbc, P1 + P2)

    # F28 (:517-557)
    Dkmab = Dklbc
    Dlmac = Dklbc
    Q1 = ein("amcl,lcbk->ambklc", VWOWO, 2.0 * blk(O, W, W, O)) - ein("amcl,lkbc->ambklc", VWOWO, blk(O, O, W, W))
    Q2 = ein("alcm,lkbc->ambklc", VWOWO, 2.0 * blk(O, O, W, W)) - ein("alcm,lcbk->ambklc", VWOWO, blk(O, W, W, O))
    t["F28"] = -ein("mb,kmab,lmac,ambklc->ak", d_ob(d, O, W), 1.0 / Dkmab, 1.0 / Dlmac, Q1 + Q2)

    # F29 (:561-606)
    Dklcd = Dklbc
    diag = ein("acbc,ckcl,klc->abkl", VWWWW, blk(W, O, W, O), 1.0 / dgv(Dklcd))
    off = (ein("cd,klcd,acbd,ckdl->abkl", lv, 1.0 / Dklcd, VWWWW, 2.0 * blk(W, O, W, O))
           - ein("cd,klcd,acbd,cldk->abkl", lv, 1.0 / Dklcd, VWWWW, blk(W, O, W, O))
           + ein("cd,klcd,adbc,dkcl->abkl", lv, 1.0 / Dklcd, VWWWW, 2.0 * blk(W, O, W, O))
           - ein("cd,klcd,adbc,dlck->abkl", lv, 1.0 / Dklcd, VWWWW, blk(W, O, W, O)))
    t["F29"] = ein("lb,klab,abkl->ak", d_ob(d, O, W), 1.0 / Dklab, diag + off)

    # F210 (:610-656)
    Dmnab = Dklbc
    dgo = lambda x: np.einsum("mmab->mab", x)
    diag = ein("ambm,mkml,mab->abkl", VWOWO, VOOOO, 1.0 / dgo(Dmnab))
    off = (ein("mn,mnab,ambn,mknl->abkl", lo, 1.0 / Dmnab, VWOWO, 2.0 * VOOOO)
           - ein("mn,mnab,ambn,mlnk->abkl", lo, 1.0 / Dmnab, VWOWO, VOOOO)
           + ein("mn,mnab,anbm,nkml->abkl", lo, 1.0 / Dmnab, VWOWO, 2.0 * VOOOO)
           - ein("mn,mnab,anbm,nlmk->abkl", lo, 1.0 / Dmnab, VWOWO, VOOOO))
    t["F210"] = ein("lb,klab,abkl->ak", d_ob(d, O, W), 1.0 / Dklab, diag + off)

    # 2h2p functions (:662-846), as full arrays over (a, b, k, l)
    ee = lambda i, j, a, b: eo[i] + eo[j] - ev[a] - ev[b]
    vv = lambda p, q, r, s: V[p, q, r, s]
    FI = np.zeros((v, o))
    FII = np.zeros((v, v, o))
    FIII = np.zeros((v, o, o))
    FIV1 = np.zeros((v, v, o, o))
    FIV2 = np.zeros((v, v, o, o))
    for a in range(v):
        A_ = o + a
        for k in range(o):
            s_ = 0.0
            for m in range(o):
                s_ -= d[m, k] / ee(m, k, a, a) * vv(A_, k, A_, m)
            for c in range(v):
                s_ += d[A_, o + c] / ee(k, k, a, c) * vv(o + c, k, A_, k)
            FI[a, k] = 2.0 * s_
    for a in range(v):
        for b in range(v):
            A_, B_ = o + a, o + b
            for k in range(o):
                s_ = 0.0
                for m in range(o):
                    s_ -= d[m, k] / ee(m, k, a, b) * (vv(A_, m, B_, k) + vv(A_, k, B_, m))
                for c in range(v):
                    s_ += d[A_, o + c] / ee(k, k, b, c) * vv(o + c, k, B_, k)
                    s_ += d[B_, o + c] / ee(k, k, a, c) * vv(o + c, k, A_, k)
                FII[a, b, k] = np.sqrt(2.0) * s_
                for l in range(o):
                    s1 = s2 = 0.0
                    for m in range(o):
                        s1 += d[m, k] / ee(m, l, a, b) * (vv(A_, m, B_, l) + vv(A_, l, m, B_))
                        s1 += d[m, l] / ee(k, m, a, b) * (vv(A_, k, m, B_) + vv(A_, m, B_, k))
                        s2 += d[m, k] / ee(m, l, a, b) * (vv(A_, m, B_, l) - vv(A_, l, m, B_))
                        s2 += d[m, l] / ee(k, m, a, b) * (vv(A_, k, m, B_) - vv(A_, m, B_, k))
                    for c in range(v):
                        C_ = o + c
                        s1 -= d[A_, C_] / ee(k, l, b, c) * (vv(C_, k, B_, l) + vv(C_, l, B_, k))
                        s1 -= d[B_, C_] / ee(k, l, a, c) * (vv(C_, l, A_, k) + vv(C_, k, A_, l))
                        s2 -= d[A_, C_] / ee(k, l, b, c) * (vv(C_, k, B_, l) - vv(C_, l, B_, k))
                        s2 -= d[B_, C_] / ee(k, l, a, c) * (vv(C_, l, A_, k) - vv(C_, k, A_, l))
                    FIV1[a, b, k, l] = s1
                    FIV2[a, b, k, l] = np.sqrt(3.0) * s2
    for a in range(v):
        A_ = o + a
        for k in range(o):
            for l in range(o):
                s_ = 0.0
                for m in range(o):
                    s_ += d[m, k] / ee(m, l, a, a) * vv(A_, m, A_, l)
                    s_ += d[m, l] / ee(k, m, a, a) * vv(A_, k, A_, m)
                for c in range(v):
                    s_ -= d[A_, o + c] / ee(k, l, a, c) * (vv(o + c, k, A_, l) + vv(o + c, l, A_, k))
                FIII[a, k, l] = np.sqrt(2.0) * s_
    t["FI"], t["FII"], t["FIII"], t["FIV1"], t["FIV2"] = FI, FII, FIII, FIV1, FIV2
    return t


def d_ob(d, O, W):
    """dpl(l, b) for occupied l, virtual b: [l, b]."""
    return d[np.ix_(O, W)]


ONE_H1P = ("F0", "F25", "F26", "F27", "F28", "F29", "F210", "FA", "FB", "FC", "F21", "F22", "F23", "F24")


def assemble(sys, dpl, terms) -> np.ndarray:
    """f (ndim) in the configuration order of oracle/csf.py."""
    o, v = sys.nocc, sys.nvirt
    d = np.asarray(dpl, dtype=np.float64)
    s = d[o:, :o].copy()
    for name in ONE_H1P:
        s = s + terms[name]
    f1 = -np.sqrt(2.0) * s                    # [a, k]
    typ, a, b, k, l = csf.table(o, v)
    f = np.zeros(typ.size)
    m0 = typ == 0
    f[m0] = f1[a[m0], k[m0]]
    m = typ == 1
    f[m] = terms["FI"][a[m], k[m]]
    m = typ == 2
    f[m] = terms["FII"][a[m], b[m], k[m]]
    m = typ == 3
    f[m] = terms["FIII"][a[m], k[m], l[m]]
    m = typ == 4
    f[m] = terms["FIV1"][a[m], b[m], k[m], l[m]]
    m = typ == 5
    f[m] = terms["FIV2"][a[m], b[m], k[m], l[m]]
    return f


def fvector(sys, dpl) -> np.ndarray:
    return assemble(sys, dpl, literal_terms(sys, dpl))


# ---------------------------------------------------------------- compact
def compact_terms(sys, dpl):
    """The same terms from the amplitudes (folded sums)."""
    o, v, eo, ev, V, d = _setup(sys, dpl)
    O, W = np.arange(o), o + np.arange(v)
    blk = lambda *ix: V[np.ix_(*ix)]
    ein = lambda s, *ops: np.einsum(s, *ops, optimize=True)
    t = mp2.amplitudes(sys, V)                # [i,j,a,b]
    U = 2.0 * t - t.transpose(0, 1, 3, 2)     # 2 t[i,j,a,b] - t[i,j,b,a]
    d_vo, d_ov, d_oo, d_vv = d[np.ix_(W, O)], d[np.ix_(O, W)], d[np.ix_(O, O)], d[np.ix_(W, W)]
    ovvv = blk(O, W, W, W)                    # (l b | a c) [l,b,a,c]
    ooov = blk(O, O, O, W)                    # (l k | m c)
    ovov = blk(O, W, O, W)                    # (k a | l b)
    oovv = blk(O, O, W, W)                    # (l k | a b)
    oooo = blk(O, O, O, O)
    vvvv = blk(W, W, W, W)
    ema = eo[None, :] - ev[:, None]           # [a,m] = e_m - e_a
    D = eo[:, None, None, None] + eo[None, :, None, None] - ev[None, None, :, None] - ev[None, None, None, :]
    r = {}
    r["F0"] = ein("bl,klab->ak", d_vo, 2.0 * t - t.transpose(1, 0, 2, 3))
    r["tauA"] = -0.5 * ein("lmab,lmcb->ca", t, 2.0 * t - t.transpose(0, 1, 3, 2))
    r["FA"] = ein("ca,ck->ak", r["tauA"], d_vo)
    r["tauB"] = -0.5 * ein("klbc,mlbc->mk", t, U)
    r["FB"] = ein("al,lk->ak", d_vo, r["tauB"])
    r["tau21"] = -ein("lmbc,lbac->am", t, 2.0 * ovvv - ovvv.transpose(0, 3, 2, 1)) / ema
    r["F21"] = ein("al,lk->ak", r["tau21"], d_oo)
    r["tau22"] = ein("lbmn,lmba->an", blk(O, W, O, O), 2.0 * t - t.transpose(1, 0, 2, 3)) / ema
    r["F22"] = ein("al,lk->ak", r["tau22"], d_oo)
    # (2 (lk|mc) - (lc|mk)) as [l,k,m,c]
    X23 = 2.0 * ooov - blk(O, W, O, O).transpose(0, 3, 2, 1)
    r["tau23"] = -ein("lmbc,lkmc->bk", t, X23) / ema
    r["F23"] = ein("ac,ck->ak", d_vv, r["tau23"])
    # (2 (db|lc) - (dc|lb)) as [d,b,l,c]
    X24 = 2.0 * blk(W, W, O, W) - blk(W, W, O, W).transpose(0, 3, 2, 1)
    r["tau24"] = ein("klbc,dblc->dk", t, X24) / ema
    r["F24"] = ein("ac,ck->ak", d_vv, r["tau24"])
    r["tauC"] = 0.5 * ein("cm,mlcb->bl", d_vo, U)
    r["FC"] = ein("bl,klab->ak", r["tauC"], 2.0 * t - t.transpose(1, 0, 2, 3))
    # K[a,k,b,l] = 2 (ak|bl) - (ab|kl); Y[a,k,b,l] = 2 t[k,l,a,b] - t[k,l,b,a]
    K = 2.0 * ovov.transpose(1, 0, 3, 2) - oovv.transpose(2, 0, 3, 1)
    Y = U.transpose(2, 0, 3, 1)
    Q = ein("akbl,blcm->akcm", K, Y) + ein("akbl,blcm->akcm", Y, K)
    r["F25+F26"] = ein("mc,kmac,akcm->ak", d_ov, 1.0 / D, Q)
    # F27: Q27[a,l,b,k] = sum_{c,m} (al|mc) Y[b,k,c,m] + (ac|ml) Y[c,k,b,m]
    Q27 = ein("lamc,bkcm->albk", ovov, Y) + ein("mlac,ckbm->albk", oovv, Y)
    r["F27"] = -ein("lb,klab,albk->ak", d_ov, 1.0 / D, Q27)
    # F28: Q28[a,m,b,k] = sum_{l,c} t[m,l,a,c] K[c,l,b,k] + t[l,m,a,c] L[c,l,b,k],
    # L[c,l,b,k] = 2 (lk|bc) - (lc|bk)
    L = 2.0 * oovv.transpose(2, 0, 3, 1) - ovov.transpose(1, 0, 3, 2)
    Q28 = ein("mlac,clbk->ambk", t, K) + ein("lmac,clbk->ambk", t, L)
    r["F28"] = -ein("mb,kmab,ambk->ak", d_ov, 1.0 / D, Q28)
    # F29: Z[a,b,k,l] = sum_{c,d} (ac|bd) Y[c,k,d,l]
    Z = ein("acbd,ckdl->abkl", vvvv, Y)
    r["F29"] = ein("lb,klab,abkl->ak", d_ov, 1.0 / D, Z)
    # F210: sum_{m,n} t[m,n,a,b] (2 (mk|nl) - (ml|nk))
    Z = ein("mnab,mknl->abkl", t, 2.0 * oooo - oooo.transpose(0, 3, 2, 1))
    r["F210"] = ein("lb,klab,abkl->ak", d_ov, 1.0 / D, Z)
    # 2h2p
    R = (ein("mk,mlab->abkl", d_oo, t) + ein("ml,kmab->abkl", d_oo, t)
         - ein("ac,klcb->abkl", d_vv, t) - ein("bc,klac->abkl", d_vv, t))
    r["R"] = R
    ia, ik = np.arange(v), np.arange(o)
    r["FI"] = -R[ia, ia][:, ik, ik]
    r["FII"] = -np.sqrt(2.0) * R[:, :, ik, ik]
    r["FIII"] = np.sqrt(2.0) * R[ia, ia]
    r["FIV1"] = R + R.transpose(1, 0, 2, 3)
    r["FIV2"] = np.sqrt(3.0) * (R - R.transpose(1, 0, 2, 3))
    return r


def fvector_compact(sys, dpl) -> np.ndarray:
    r = dict(compact_terms(sys, dpl))
    r["F25"], r["F26"] = r.pop("F25+F26"), 0.0
    return assemble(sys, dpl, r)
