"""Density-fitted restatement of the ADC(2)-s sigma-vector.  Test oracle only.

A second, independent formulation of the same matrix as ``elements``: every
coupling function C1..C5 of source/adclib/adc_ph.f90:974-1073 is a
Kronecker-delta selection of (oo|ov) and (vv|ov) integrals, so with
(pq|rs) = sum_P B[P,p,q] B[P,r,s] the whole 1h1p<->2h2p coupling becomes

  sigma1(a,k) = sum_{P,c} B[P,a,c] Y[P,c,k] - sum_{P,m} B[P,k,m] Y[P,a,m],
  Y[P,c,k]    = sum_{b,l} B[P,l,b] Z(c,b,k,l),
  sigma2(J)   = projection onto CSF J of G(c,b,k,l) = sum_P W[P,c,k] B[P,l,b],
  W[P,c,k]    = sum_a B[P,c,a] x1(a,k) - sum_m B[P,k,m] x1(c,m),

where Z is the closed-shell pair tensor assembled from the five CSF blocks
(``pair_tensor``) and ``csf_from_pairs`` is its transpose.  The ph-ph block
(adc_ph.f90:31-53, 769-957) is rebuilt from the same factors by the
separable-denominator products in ``phph_block``.

Pinned by: elementwise agreement with the dense restatement of ``elements``
(test_oracle_pins), which evaluates the reference's element functions
literally; the two share no code beyond the configuration table.
"""
from __future__ import annotations

import numpy as np

from . import csf

R2 = np.sqrt(2.0)
S4 = 1.0 / np.sqrt(2.0)
S5 = np.sqrt(1.5)


def _pairs(n):
    i, j = np.triu_indices(n, 1)
    return i, j


def pair_tensor(x2: np.ndarray, o: int, v: int) -> np.ndarray:
    """CSF 2h2p amplitudes (n2,) -> Z[c,b,k,l] (v,v,o,o)."""
    c = csf.counts(o, v)
    Z = np.zeros((v, v, o, o))
    off = 0
    x = x2[off:off + c["t1"]].reshape(o, v)          # t1: [i][a]
    off += c["t1"]
    ar = np.arange(v)
    orr = np.arange(o)
    Z[ar[None, :], ar[None, :], orr[:, None], orr[:, None]] = -R2 * x
    pa, pb = _pairs(v)
    x = x2[off:off + c["t2"]].reshape(pa.size, o)     # t2: [(a<b)][i]
    off += c["t2"]
    Z[pa[:, None], pb[:, None], orr[None, :], orr[None, :]] = -x
    Z[pb[:, None], pa[:, None], orr[None, :], orr[None, :]] = -x
    pk, pl = _pairs(o)
    x = x2[off:off + c["t3"]].reshape(pk.size, v)     # t3: [(i<j)][a]
    off += c["t3"]
    Z[ar[None, :], ar[None, :], pk[:, None], pl[:, None]] = x
    Z[ar[None, :], ar[None, :], pl[:, None], pk[:, None]] = x
    x4 = x2[off:off + c["t4"]].reshape(pk.size, pa.size)  # t4: [(i<j)][(a<b)]
    off += c["t4"]
    x5 = x2[off:off + c["t5"]].reshape(pk.size, pa.size)
    s4 = S4 * x4
    s5 = S5 * x5
    A, Bb = pa[None, :], pb[None, :]
    K, L = pk[:, None], pl[:, None]
    Z[A, Bb, K, L] = s4 + s5
    Z[Bb, A, L, K] = s4 + s5
    Z[Bb, A, K, L] = s4 - s5
    Z[A, Bb, L, K] = s4 - s5
    return Z


def csf_from_pairs(G: np.ndarray, o: int, v: int) -> np.ndarray:
    """Transpose of ``pair_tensor``: G[c,b,k,l] -> CSF 2h2p vector (n2,)."""
    ar = np.arange(v)
    orr = np.arange(o)
    pa, pb = _pairs(v)
    pk, pl = _pairs(o)
    out = []
    out.append((-R2 * G[ar[None, :], ar[None, :], orr[:, None], orr[:, None]]).ravel())
    out.append((-(G[pa[:, None], pb[:, None], orr[None, :], orr[None, :]]
                  + G[pb[:, None], pa[:, None], orr[None, :], orr[None, :]])).ravel())
    out.append((G[ar[None, :], ar[None, :], pk[:, None], pl[:, None]]
                + G[ar[None, :], ar[None, :], pl[:, None], pk[:, None]]).ravel())
    A, Bb = pa[None, :], pb[None, :]
    K, L = pk[:, None], pl[:, None]
    g1 = G[A, Bb, K, L]
    g2 = G[Bb, A, K, L]
    g3 = G[A, Bb, L, K]
    g4 = G[Bb, A, L, K]
    out.append((S4 * (g1 + g2 + g3 + g4)).ravel())
    out.append((S5 * (g1 - g2 - g3 + g4)).ravel())
    return np.concatenate(out)


def diag2(sys) -> np.ndarray:
    """2h2p diagonal e_a + e_b - e_k - e_l in CSF order."""
    o, v = sys.nocc, sys.nvirt
    e = sys.eps
    eo, ev = e[:o], e[o:]
    pa, pb = _pairs(v)
    pk, pl = _pairs(o)
    d = [(ev[None, :] * 2 - 2 * eo[:, None]).ravel(),
         (ev[pa][:, None] + ev[pb][:, None] - 2 * eo[None, :]).ravel(),
         (2 * ev[None, :] - eo[pk][:, None] - eo[pl][:, None]).ravel()]
    d4 = (ev[pa][None, :] + ev[pb][None, :] - eo[pk][:, None] - eo[pl][:, None]).ravel()
    d += [d4, d4]
    return np.concatenate(d)


class DF:
    def __init__(self, sys):
        o, v = sys.nocc, sys.nvirt
        self.o, self.v, self.sys = o, v, sys
        B = sys.B
        self.Boo = B[:, :o, :o]
        self.Bvv = B[:, o:, o:]
        self.Bov = np.ascontiguousarray(B[:, o:, :o]).reshape(B.shape[0], v * o)  # [P,(a,k)]
        self.n1 = o * v
        self.D2 = diag2(sys)
        self.H11 = phph_block(sys)

    def couple_down(self, x2):
        o, v = self.o, self.v
        Z = pair_tensor(x2, o, v)                                    # [c,b,k,l]
        Zm = Z.transpose(0, 2, 1, 3).reshape(self.n1, self.n1)      # [(c,k),(b,l)]
        Y = (self.Bov @ Zm.T).reshape(-1, v, o)                     # Y[P,c,k]
        s = (np.einsum("Pac,Pck->ak", self.Bvv, Y, optimize=True)
             - np.einsum("Pkm,Pam->ak", self.Boo, Y, optimize=True))
        return s.reshape(self.n1)

    def couple_up(self, x1):
        o, v = self.o, self.v
        X = x1.reshape(v, o)
        W = (np.einsum("Pca,ak->Pck", self.Bvv, X, optimize=True)
             - np.einsum("Pkm,cm->Pck", self.Boo, X, optimize=True))
        Gm = W.reshape(-1, self.n1).T @ self.Bov                    # [(c,k),(b,l)]
        G = Gm.reshape(v, o, v, o).transpose(0, 2, 1, 3)            # [c,b,k,l]
        return csf_from_pairs(G, o, v)

    def sigma(self, X):
        X = np.asarray(X, dtype=np.float64)
        one = X.ndim == 1
        if one:
            X = X[:, None]
        n1 = self.n1
        out = np.empty_like(X)
        for j in range(X.shape[1]):
            x1 = X[:n1, j]
            x2 = X[n1:, j]
            out[:n1, j] = self.H11 @ x1 + self.couple_down(x2)
            out[n1:, j] = self.couple_up(x1) + self.D2 * x2
        return out[:, 0] if one else out

    def diag(self):
        return np.concatenate([np.diag(self.H11), self.D2])


def phph_block(sys) -> np.ndarray:
    """ph-ph block (n1 x n1, rows (a,k) a-major) from the DF factors.

    First order: (e_a - e_k) delta + 2 (ak|a'k') - (aa'|kk').
    Second order (with D_a(r;u,v) = e_r + e_a - e_u - e_v, K(a;u,v,r) = (au|rv)):
      CA(a,a') = 1/2 sum_{ruv} (1/D_a + 1/D_a') K(a;uvr) [2K(a';uvr) - K(a';vur)]
      CB(k,k') = -1/2 sum_{cdi} (1/(e_i-e_c-e_d+e_k) + 1/(..+e_k'))
                       (ck|di) [2(ck'|di) - (dk'|ci)]
      CC       = -1/2 [(T/D) T^T + T (T/D)^T],  T(ak,ci) = 2(ak|ci) - (ai|ck),
                 D(ak,ci) = e_a + e_c - e_k - e_i
    """
    o, v = sys.nocc, sys.nvirt
    e = sys.eps
    eo, ev = e[:o], e[o:]
    B = sys.B
    n1 = o * v
    Bov = np.ascontiguousarray(B[:, o:, :o]).reshape(B.shape[0], n1)
    M = (Bov.T @ Bov).reshape(v, o, v, o)               # M[a,k,c,i] = (ak|ci)
    J = np.einsum("Pab,Pkl->akbl", B[:, o:, o:], B[:, :o, :o], optimize=True)
    H = (2.0 * M - J).reshape(n1, n1)
    H[np.arange(n1), np.arange(n1)] += (ev[:, None] - eo[None, :]).ravel()
    # CA: K(a;u,v,r) = M[a,u,r,v]
    Kt = M.transpose(0, 1, 3, 2)                        # [a,u,v,r]
    Lt = 2.0 * Kt - Kt.transpose(0, 2, 1, 3)            # [a',u,v,r]: 2K(a';uvr) - K(a';vur)
    Dr = ev[None, None, None, :] + ev[:, None, None, None] - eo[None, :, None, None] - eo[None, None, :, None]
    Kd = (Kt / Dr).reshape(v, -1)
    Kf = Kt.reshape(v, -1)
    Lf = Lt.reshape(v, -1)
    Ld = (Lt / Dr).reshape(v, -1)
    CA = 0.5 * (Kd @ Lf.T + Kf @ Ld.T)
    # CB: P(k; c,d,i) = M[c,k,d,i], Q(k'; c,d,i) = 2 M[c,k',d,i] - M[d,k',c,i]
    Pk = M.transpose(1, 0, 2, 3)                        # [k,c,d,i]
    Qk = 2.0 * Pk - Pk.transpose(0, 2, 1, 3)
    Ed = eo[None, None, None, :] - ev[None, :, None, None] - ev[None, None, :, None] + eo[:, None, None, None]
    Pd = (Pk / Ed).reshape(o, -1)
    Pf = Pk.reshape(o, -1)
    Qf = Qk.reshape(o, -1)
    Qd = (Qk / Ed).reshape(o, -1)
    CB = -0.5 * (Pd @ Qf.T + Pf @ Qd.T)
    # CC
    T = (2.0 * M - M.transpose(0, 3, 2, 1)).reshape(n1, n1)
    Dd = (ev[:, None] - eo[None, :]).ravel()
    Dm = Dd[:, None] + Dd[None, :]
    TD = T / Dm
    CC = -0.5 * (TD @ T.T + T @ TD.T)
    ii = np.arange(v)
    kk = np.arange(o)
    H = H.reshape(v, o, v, o)
    H[:, kk, :, kk] += CA[None, :, :]
    H[ii, :, ii, :] += CB[None, :, :]
    return H.reshape(n1, n1) + CC
