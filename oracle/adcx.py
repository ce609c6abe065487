"""ADC(2)-x: the first-order 2h2p-2h2p block.  Test oracle only.

The reference's ADC(2)-x matrix (get_offdiag_adc2ext_save,
source/adclib/get_matrix.f90:2435-2893, with the element functions C_1_1 ..
C_4ii_4ii of source/adclib/adc_ph.f90:1102-1645, and the diagonal of
get_diag_adc2ext_direct, get_matrix.f90) is the ADC(2)-s matrix with the
first-order 2h2p-2h2p block M22 added to the diagonal 2h2p block.

Restated here in pair space: with Z = pair_tensor(Nt x2) (oracle/dfsigma.py:
the closed-shell pair tensor of the five CSF blocks) and Nt = 1 on t1..t4
and 1/3 on t5 (the ratio of the norms of the CSF states that pair_tensor
maps to: sum_cbkl Z(c,b,k,l) E_ck E_bl |0>),
    T = 2 Z - X Z,        (X Z)(c,b,k,l) = Z(b,c,k,l),
    G(c,b,k,l) = 1/2 sum_de (cd|be) T(d,e,k,l) + 1/2 sum_mn (km|ln) T(c,b,m,n)
               + sum_dm [2 (ck|dm) - (cd|km)] T(d,b,m,l)
               - sum_dm [(cl|dm) T(d,b,m,k) + (cd|lm) T(d,b,k,m)],
    M22 x2 = Nt csf_from_pairs(G).
This form was found by fitting pair-space contractions to the block itself
and is pinned by tests/test_oracle_pins.py against the first-order
intermediate-state matrix <Psi_J|V - E1|Psi_J'> built exactly in the
determinant space on the same CSF states (tests/detspace.py), element by
element, which is what the reference's element functions are derived to be;
the element functions C_1_1, C_2_2, C_3_3 and C_4ii_4ii (the diagonal the
reference adds to K_2p2h_2p2h) and C_4ii_1 and C_3_2 (two off-diagonal blocks
with their assembly order) are restated literally below and agree with both.
"""
from __future__ import annotations

import numpy as np

from . import csf, dfsigma, mp2


def _nt(o: int, v: int) -> np.ndarray:
    c = csf.counts(o, v)
    nt = np.ones(c["n2"])
    nt[c["n2"] - c["t5"]:] = 1.0 / 3.0
    return nt


def pair_operator(eri: np.ndarray, o: int, v: int, Z: np.ndarray) -> np.ndarray:
    """G of the module docstring for a pair tensor Z[c,b,k,l]."""
    O, W = slice(0, o), slice(o, o + v)
    vvvv = eri[W, W, W, W]            # (c d | b e) -> [c,d,b,e]
    oooo = eri[O, O, O, O]            # (k m | l n) -> [k,m,l,n]
    ovvo = eri[W, O, W, O]            # (c k | d m) -> [c,k,d,m]
    vvoo = eri[W, W, O, O]            # (c d | k m) -> [c,d,k,m]
    T = 2.0 * Z - Z.transpose(1, 0, 2, 3)
    ein = lambda s, *a: np.einsum(s, *a, optimize=True)  # noqa: E731
    G = 0.5 * ein("cdbe,dekl->cbkl", vvvv, T) + 0.5 * ein("kmln,cbmn->cbkl", oooo, T)
    G += ein("ckdm,dbml->cbkl", 2.0 * ovvo - vvoo.transpose(0, 2, 1, 3), T)
    G -= ein("cldm,dbmk->cbkl", ovvo, T) + ein("cdlm,dbkm->cbkl", vvoo, T)
    return G


def block22(sys, x2: np.ndarray) -> np.ndarray:
    """M22 x2 for a 2h2p vector (n2,) or block (n2, nvec)."""
    o, v = sys.nocc, sys.nvirt
    eri = mp2.integrals(sys)
    nt = _nt(o, v)
    X = x2.reshape(x2.shape[0], -1)
    out = np.empty_like(X)
    for j in range(X.shape[1]):
        Z = dfsigma.pair_tensor(nt * X[:, j], o, v)
        out[:, j] = nt * dfsigma.csf_from_pairs(pair_operator(eri, o, v, Z), o, v)
    return out.reshape(x2.shape)


def block22_dense(sys) -> np.ndarray:
    n2 = csf.counts(sys.nocc, sys.nvirt)["n2"]
    return block22(sys, np.eye(n2))


# ---- literal restatement of four element functions (adc_ph.f90), as the
# reference's ADC(2)-x diagonal uses them (get_diag_adc2ext_direct); p, q, r,
# s are absolute orbital indices, vpqrs(p,q,r,s) = (pq|rs)
def c_1_1(V, a, k, ap, kp):
    f = 0.0
    if k == kp:
        f += V[a, ap, a, ap]
    if a == ap:
        f += V[kp, k, kp, k]
    if a == ap and k == kp:
        f -= 2.0 * (2.0 * V[a, ap, kp, k] - V[a, k, kp, ap])
    return f


def c_2_2(V, a, b, k, ap, bp, kp):
    f = 0.0
    if k == kp:
        f += V[a, ap, b, bp] + V[a, bp, b, ap]
    if a == ap and b == bp:
        f += V[kp, k, kp, k]
    if b == bp and k == kp:
        f -= 2.0 * V[a, ap, kp, k] - V[a, k, kp, ap]
    if a == ap and k == kp:
        f -= 2.0 * V[b, bp, kp, k] - V[b, k, kp, bp]
    if b == ap and k == kp:
        f -= 2.0 * V[a, bp, kp, k] - V[a, k, kp, bp]
    if a == bp and k == kp:
        f -= 2.0 * V[b, ap, kp, k] - V[b, k, kp, ap]
    return f


def c_3_3(V, a, k, l, ap, kp, lp):
    f = 0.0
    if k == kp and l == lp:
        f += V[a, ap, a, ap]
    if a == ap:
        f += V[k, kp, lp, l] + V[kp, l, lp, k]
    if a == ap and l == lp:
        f -= 2.0 * V[a, ap, kp, k] - V[a, k, kp, ap]
    if a == ap and k == kp:
        f -= 2.0 * V[a, ap, lp, l] - V[a, l, lp, ap]
    if a == ap and l == kp:
        f -= 2.0 * V[a, ap, lp, k] - V[a, k, lp, ap]
    if a == ap and k == lp:
        f -= 2.0 * V[a, ap, kp, l] - V[a, l, kp, ap]
    return f


def c_4ii_4ii(V, a, b, k, l, ap, bp, kp, lp):
    f = 0.0
    if k == kp and l == lp:
        f += V[a, ap, b, bp] - V[a, bp, b, ap]
    if a == ap and b == bp:
        f += V[kp, k, lp, l] - V[kp, l, lp, k]
    for (x, xp, y, yp, s) in ((b, bp, a, ap, 1.0), (a, ap, b, bp, 1.0), (b, ap, a, bp, -1.0), (a, bp, b, ap, -1.0)):
        # the four virtual pairings with the sign of the term list
        if x == xp:
            if l == lp:
                f -= s * (V[y, yp, kp, k] - 1.5 * V[y, k, kp, yp])
            if k == kp:
                f -= s * (V[y, yp, lp, l] - 1.5 * V[y, l, lp, yp])
            if l == kp:
                f += s * (V[y, yp, lp, k] - 1.5 * V[y, k, lp, yp])
            if k == lp:
                f += s * (V[y, yp, kp, l] - 1.5 * V[y, l, kp, yp])
    return f


def c_4ii_1(V, a, b, k, l, ap, kp):
    f = 0.0
    if b == ap and k == kp:
        f += V[a, l, kp, ap]
    if a == ap and l == kp:
        f += V[b, k, kp, ap]
    if b == ap and l == kp:
        f -= V[a, k, kp, ap]
    if a == ap and k == kp:
        f -= V[b, l, kp, ap]
    return np.sqrt(3.0) * f


def c_3_2(V, a, k, l, ap, bp, kp):
    f = 0.0
    if k == kp and l == kp:
        f -= V[a, ap, a, bp]
    if a == ap and a == bp:
        f -= V[kp, k, kp, l]
    if a == bp and l == kp:
        f += 2.0 * V[a, ap, kp, k] - V[a, k, kp, ap]
    if a == bp and k == kp:
        f += 2.0 * V[a, ap, kp, l] - V[a, l, kp, ap]
    if a == ap and l == kp:
        f += 2.0 * V[a, bp, kp, k] - V[a, k, kp, bp]
    if a == ap and k == kp:
        f += 2.0 * V[a, bp, kp, l] - V[a, l, kp, bp]
    return f


def offdiagonal_blocks(sys):
    """Two off-diagonal blocks of the reference's ADC(2)-x 2h2p-2h2p matrix by
    its element functions and assembly (get_matrix.f90 (4ii,1) and (3,2)
    blocks): rows t5 x columns t1 with C_4ii_1, rows t3 x columns t2 with
    C_3_2; returned as {name: (rows, cols, values)} in 2h2p indices."""
    o, v = sys.nocc, sys.nvirt
    V = mp2.integrals(sys)
    typ, a, b, k, l = csf.table(o, v)
    n1 = o * v
    idx = {t: np.where(typ == t)[0] for t in range(1, 6)}
    out = {}
    r, c = idx[5], idx[1]
    out["t5,t1"] = (r - n1, c - n1, np.array([[c_4ii_1(V, o + a[i], o + b[i], k[i], l[i], o + a[j], k[j])
                                               for j in c] for i in r]))
    r, c = idx[3], idx[2]
    out["t3,t2"] = (r - n1, c - n1, np.array([[c_3_2(V, o + a[i], k[i], l[i], o + a[j], o + b[j], k[j])
                                               for j in c] for i in r]))
    return out


def diagonal_first_order(sys) -> np.ndarray:
    """The first-order part of the reference's ADC(2)-x 2h2p diagonal
    (get_diag_adc2ext_direct: K_2p2h_2p2h + C_1_1 / C_2_2 / C_3_3 / C_4i_4i /
    C_4ii_4ii of each configuration with itself), for t1, t2, t3 and t5 (the
    t4 function C_4i_4i is not restated)."""
    o, v = sys.nocc, sys.nvirt
    V = mp2.integrals(sys)
    typ, a, b, k, l = csf.table(o, v)
    n1 = o * v
    out = np.full(typ.size - n1, np.nan)
    for J in range(n1, typ.size):
        A, Bv, K, L = o + a[J], o + b[J], k[J], l[J]
        t = typ[J]
        if t == 1:
            out[J - n1] = c_1_1(V, A, K, A, K)
        elif t == 2:
            out[J - n1] = c_2_2(V, A, Bv, K, A, Bv, K)
        elif t == 3:
            out[J - n1] = c_3_3(V, A, K, L, A, K, L)
        elif t == 5:
            out[J - n1] = c_4ii_4ii(V, A, Bv, K, L, A, Bv, K, L)
    return out
