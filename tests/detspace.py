"""Brute-force determinant-space perturbation theory for tiny closed shells.
Test helper (a checker of the oracle, independent of the reference's
formulas).

The Hamiltonian of a synthetic system (orbital energies eps, integrals
(pq|rs) = sum_P B[P,p,q] B[P,r,s]) is built in second quantisation over the
Sz = 0 determinants of 2*nocc electrons in 2*nbas spin orbitals, with the
one-electron part chosen so that the Fock matrix is diag(eps):
    h_pq = eps_p delta_pq - sum_j [2 (pq|jj) - (pj|jq)],
    H = sum_{pq,s} h_pq a+_ps a_qs + 1/2 sum_{pqrs,s,t} (pq|rs) a+_ps a+_rt a_st a_qs.
With H0 = sum_p eps_p n_p (diagonal in determinants) and V = H - H0,
Rayleigh-Schroedinger perturbation theory in intermediate normalisation is
done exactly by linear algebra in the determinant space:
    Psi1 = R V Phi0,  Psi2 = R (V - E1) Psi1,  R = Q (E0 - H0)^-1 Q,
    E2 = <Phi0|V|Psi1>,
    rho2 = <Psi1|g|Psi1> + <Phi0|g|Psi2> + <Psi2|g|Phi0> - <Psi1|Psi1> <Phi0|g|Phi0>
for the spin-summed one-particle density g_pq = sum_s a+_ps a_qs.

Intermediate-state representation (ISR) of the 1h1p class, also exact in the
determinant space: for H(lam) = H0 + lam V with exact ground state Psi0(lam)
(lowest eigenvector, <Phi0|Psi0> > 0), the precursor states
C_ak Psi0 (C_ak = g_ak / sqrt(2), the singlet 1h1p excitation), made
orthogonal to Psi0 and then symmetrically (Loewdin) orthonormalised among
themselves, give the intermediate states Psi~_ak(lam), and
    f_ak(lam) = <Psi~_ak|D|Psi0>,
    M_ak,bl(lam) = <Psi~_ak|H(lam) - E0(lam)|Psi~_bl>.
Their Taylor coefficients in lam are the strict perturbation-theory orders;
they are read off a polynomial through samples at lam = h*(-n..n).
"""
from __future__ import annotations

import itertools

import numpy as np


def _sign_before(det: int, k: int) -> int:
    return -1 if bin(det & ((1 << k) - 1)).count("1") % 2 else 1


def _annihilate(det, k):
    if not (det >> k) & 1:
        return None, 0
    return det ^ (1 << k), _sign_before(det, k)


def _create(det, k):
    if (det >> k) & 1:
        return None, 0
    return det | (1 << k), _sign_before(det, k)


class DetSpace:
    """Sz = 0 determinants of nocc alpha and nocc beta electrons in nbas
    spatial orbitals; spin orbital 2p + s (s = 0 alpha, 1 beta)."""

    def __init__(self, nocc: int, nbas: int):
        self.nocc, self.nbas = nocc, nbas
        dets = []
        for a in itertools.combinations(range(nbas), nocc):
            for b in itertools.combinations(range(nbas), nocc):
                d = 0
                for p in a:
                    d |= 1 << (2 * p)
                for p in b:
                    d |= 1 << (2 * p + 1)
                dets.append(d)
        self.dets = dets
        self.index = {d: i for i, d in enumerate(dets)}
        self.ref = self.index[sum((1 << (2 * p)) | (1 << (2 * p + 1)) for p in range(nocc))]

    def one_body(self, p: int, q: int) -> np.ndarray:
        """Matrix of g_pq = sum_s a+_ps a_qs."""
        n = len(self.dets)
        M = np.zeros((n, n))
        for j, d in enumerate(self.dets):
            for s in (0, 1):
                d1, s1 = _annihilate(d, 2 * q + s)
                if d1 is None:
                    continue
                d2, s2 = _create(d1, 2 * p + s)
                if d2 is None:
                    continue
                M[self.index[d2], j] += s1 * s2
        return M

    def hamiltonian(self, h: np.ndarray, eri: np.ndarray) -> np.ndarray:
        nb = self.nbas
        n = len(self.dets)
        H = np.zeros((n, n))
        for p in range(nb):
            for q in range(nb):
                if h[p, q] != 0.0:
                    H += h[p, q] * self.one_body(p, q)
        for j, d in enumerate(self.dets):
            for q, s, r, t in itertools.product(range(nb), (0, 1), range(nb), (0, 1)):
                # a+_ps a+_rt a_st a_qs: annihilate q_s then s_t, create r_t then p_s
                d1, c1 = _annihilate(d, 2 * q + s)
                if d1 is None:
                    continue
                for ss in range(nb):
                    d2, c2 = _annihilate(d1, 2 * ss + t)
                    if d2 is None:
                        continue
                    d3, c3 = _create(d2, 2 * r + t)
                    if d3 is None:
                        continue
                    for p in range(nb):
                        val = eri[p, q, r, ss]
                        if val == 0.0:
                            continue
                        d4, c4 = _create(d3, 2 * p + s)
                        if d4 is None:
                            continue
                        H[self.index[d4], j] += 0.5 * val * c1 * c2 * c3 * c4
        return H


def mp_second_order(eps: np.ndarray, eri: np.ndarray, nocc: int):
    """(E2, rho0, rho2) of the closed shell by exact RSPT in the determinant space."""
    nb = eps.size
    o = nocc
    h = np.diag(eps).astype(np.float64)
    h -= 2.0 * np.einsum("pqjj->pq", eri[:, :, :o, :o]) - np.einsum("pjjq->pq", eri[:, :o, :o, :])
    ds = DetSpace(o, nb)
    H = ds.hamiltonian(h, eri)
    occ_energy = np.array([sum(eps[k // 2] for k in range(2 * nb) if (d >> k) & 1) for d in ds.dets])
    E0 = occ_energy[ds.ref]
    V = H - np.diag(occ_energy)
    n = len(ds.dets)
    R = np.zeros(n)
    mask = np.arange(n) != ds.ref
    R[mask] = 1.0 / (E0 - occ_energy[mask])
    phi0 = np.zeros(n)
    phi0[ds.ref] = 1.0
    psi1 = R * (V @ phi0)
    E1 = phi0 @ V @ phi0
    psi2 = R * (V @ psi1 - E1 * psi1)
    E2 = phi0 @ V @ psi1
    rho0 = np.zeros((nb, nb))
    rho2 = np.zeros((nb, nb))
    n11 = psi1 @ psi1
    for p in range(nb):
        for q in range(nb):
            g = ds.one_body(p, q)
            rho0[p, q] = phi0 @ g @ phi0
            rho2[p, q] = psi1 @ g @ psi1 + phi0 @ g @ psi2 + psi2 @ g @ phi0 - n11 * rho0[p, q]
    return E2, rho0, rho2


def _fock_hamiltonian(eps: np.ndarray, eri: np.ndarray, nocc: int):
    """(DetSpace, H, diagonal of H0): the Hamiltonian whose Fock matrix is diag(eps)."""
    nb = eps.size
    o = nocc
    h = np.diag(eps).astype(np.float64)
    h -= 2.0 * np.einsum("pqjj->pq", eri[:, :, :o, :o]) - np.einsum("pjjq->pq", eri[:, :o, :o, :])
    ds = DetSpace(o, nb)
    H = ds.hamiltonian(h, eri)
    occ_energy = np.array([sum(eps[k // 2] for k in range(2 * nb) if (d >> k) & 1) for d in ds.dets])
    return ds, H, occ_energy


def isr_orders(eps: np.ndarray, eri: np.ndarray, nocc: int, dpl: np.ndarray, h: float = 0.02, npts: int = 5):
    """Strict perturbation orders of the ISR quantities of the ADC(2) classes.

    1h1p class (singlet C_ak = g_ak / sqrt(2), (a, k) virtual-major a*nocc + k):
      f1[n, a, k] and M11[n] (orders n = 0, 1, 2).
    2h2p class: precursors g_ak g_bl Psi0 for unordered pairs (a,k) <= (b,l)
    (they span the singlet 2h2p space), made orthogonal to Psi0 and to the
    1h1p intermediate states, then Loewdin-orthonormalised.  This basis
    differs from any other orthonormal singlet 2h2p basis by an orthogonal
    transformation, so only invariants under it are returned, each as its
    lam^0, lam^1, lam^2 coefficients (the first two vanish):
      CC  = C^T C,  CDC = C^T M22 C  (1h1p x 1h1p),  Cf = C^T f2,  ff = f2^T f2,
    with C = <Psi~_2h2p|H - E0|Psi~_1h1p> and f2 = <Psi~_2h2p|D|Psi0>; their
    lam^2 coefficients are C1^T C1, C1^T D0 C1, C1^T f2^(1), |f2^(1)|^2 of the
    first-order coupling C1, zeroth-order 2h2p diagonal D0 and first-order
    2h2p f-vector f2^(1)."""
    nb = eps.size
    o, v = nocc, nb - nocc
    n1 = o * v
    ds, H, e0 = _fock_hamiltonian(eps, eri, nocc)
    V = H - np.diag(e0)
    D = np.zeros_like(H)
    G = {}
    for p in range(nb):
        for q in range(nb):
            g = ds.one_body(p, q)
            D += dpl[p, q] * g
            G[p, q] = g
    C1 = [G[o + a, k] / np.sqrt(2.0) for a in range(v) for k in range(o)]
    E = [G[o + a, k] for a in range(v) for k in range(o)]
    C2 = [E[i] @ E[j] for i in range(n1) for j in range(i, n1)]
    lams = h * np.arange(-npts, npts + 1)
    rec = {k: [] for k in ("f1", "M11", "CC", "CDC", "Cf", "ff")}
    for lam in lams:
        Hl = np.diag(e0) + lam * V
        w, U = np.linalg.eigh(Hl)
        psi = U[:, 0] * np.sign(U[ds.ref, 0])
        Hs = Hl - w[0] * np.eye(Hl.shape[0])

        def loewdin(P):
            sw, sv = np.linalg.eigh(P.T @ P)
            return P @ (sv @ np.diag(sw ** -0.5) @ sv.T)

        P1 = np.array([g @ psi for g in C1]).T
        P1 = loewdin(P1 - np.outer(psi, psi @ P1))
        P2 = np.array([g @ psi for g in C2]).T
        P2 = P2 - np.outer(psi, psi @ P2) - P1 @ (P1.T @ P2)
        P2 = loewdin(P2)
        Cm = P2.T @ Hs @ P1
        f2 = P2.T @ D @ psi
        rec["f1"].append(P1.T @ D @ psi)
        rec["M11"].append(P1.T @ Hs @ P1)
        rec["CC"].append(Cm.T @ Cm)
        rec["CDC"].append(Cm.T @ (P2.T @ Hs @ P2) @ Cm)
        rec["Cf"].append(Cm.T @ f2)
        rec["ff"].append(np.array([f2 @ f2]))
    Vm = np.vander(lams, lams.size, increasing=True)
    out = {}
    for k, vals in rec.items():
        arr = np.array(vals)
        c = np.linalg.solve(Vm, arr.reshape(lams.size, -1))[:3]
        out[k] = c.reshape((3,) + arr.shape[1:])
    out["f1"] = out["f1"].reshape(3, v, o)
    return out


def csf_first_order(eps: np.ndarray, eri: np.ndarray, nocc: int, pair_tensor):
    """(C1, M22): first-order coupling <Phi_ak|V|Phi_J> and 2h2p block
    <Phi_J|V - E1|Phi_J'> between the determinant-space states of the
    singlet CSFs: Phi_ak = g_ak/sqrt(2) Phi0 and, for a 2h2p CSF J,
    Phi_J proportional to sum_cbkl Z_J(c,b,k,l) g_ck g_bl Phi0 with
    Z_J = pair_tensor(e_J), normalised.  At first order the intermediate
    states are these states, so C1 and M22 are the first-order ISR blocks."""
    nb = eps.size
    o, v = nocc, nb - nocc
    ds, H, e0 = _fock_hamiltonian(eps, eri, nocc)
    V = H - np.diag(e0)
    g = {(p, q): ds.one_body(p, q) for p in range(nb) for q in range(nb)}
    phi0 = np.zeros(len(ds.dets))
    phi0[ds.ref] = 1.0
    P1 = np.array([g[o + a, k] @ phi0 / np.sqrt(2.0) for a in range(v) for k in range(o)]).T
    cols = []
    n2 = _n2(o, v)
    for J in range(n2):
        e = np.zeros(n2)
        e[J] = 1.0
        Z = pair_tensor(e, o, v)
        s = np.zeros_like(phi0)
        for c, b, k, l in zip(*np.nonzero(Z)):
            s += Z[c, b, k, l] * (g[o + c, k] @ (g[o + b, l] @ phi0))
        cols.append(s / np.linalg.norm(s))
    P2 = np.array(cols).T
    E1 = phi0 @ V @ phi0
    return P1.T @ V @ P2, P2.T @ V @ P2 - E1 * np.eye(n2), P2.T @ P2


def _n2(o: int, v: int) -> int:
    npv, npo = v * (v - 1) // 2, o * (o - 1) // 2
    return o * v + npv * o + npo * v + 2 * npo * npv
