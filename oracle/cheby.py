"""Chebyshev order-domain autocorrelation function.  Test oracle only.

Restates chebyshev_auto (source/propagation/chebyspec.f90:245-319) with the
scaled matrix-vector product of matxvec_chebyshev_incore
(source/propagation/tdselib.f90:1316-1398):

    H_norm = (2/dE) H - (2 Emin/dE + 1) I,   dE = Emax - Emin,
    q_1 = H_norm q_0,   q_k = 2 H_norm q_{k-1} - q_{k-2}     (chebyshev_recursion,
                                                             tdselib.f90:1228-1257)
    C_k = q_0 . q_k                                  k = 0..K
    C_2k = 2 q_k . q_k - C_0,  C_{2k-1} = 2 q_{k-1} . q_k - C_1   for k > K/2 (integer division)

Pinned by the closed form C_k = sum_n (v_n . q_0)^2 T_k(x_n), x_n the
eigenvalues of H_norm (``closed_form``, dense eigendecomposition), which
tests/test_oracle_pins.py checks on small systems.

``fdstates`` restates calc_eigenstates of fdstates_cfd
(source/propagation/fdstates_cfd.f90:207-310): v_i = sum_k coef[k, i] q_k with
the same recursion, normalised, E_i = v_i . H v_i (matxvec returns -H v, which
the reference negates).  Pinned by T_k(H_norm) q0 = V cos(k arccos x) V^T q0
from the dense eigendecomposition, and by coefficients that make a filter
which keeps one eigenvector (the state and its eigenvalue come back).
"""
from __future__ import annotations

import numpy as np


def chebyshev_auto(matvec, q0: np.ndarray, bounds, order: int) -> np.ndarray:
    """C_0..C_{2K} for K = order; ``matvec(x)`` returns H x."""
    emin, emax = float(bounds[0]), float(bounds[1])
    de = emax - emin
    a, b = 2.0 / de, -(2.0 * emin / de + 1.0)
    auto = np.zeros(2 * order + 1)
    q0 = np.asarray(q0, dtype=np.float64)
    auto[0] = q0 @ q0
    qkm1, qkm2 = q0.copy(), None
    for k in range(1, order + 1):
        qk = a * matvec(qkm1) + b * qkm1
        if k > 1:
            qk = 2.0 * qk - qkm2
        auto[k] = q0 @ qk
        if k > order // 2:
            auto[2 * k] = 2.0 * (qk @ qk) - auto[0]
            auto[2 * k - 1] = 2.0 * (qkm1 @ qk) - auto[1]
        qkm2, qkm1 = qkm1, qk
    return auto


def closed_form(H: np.ndarray, q0: np.ndarray, bounds, nk: int) -> np.ndarray:
    """sum_n (v_n . q0)^2 T_k(x_n) for k = 0..nk-1 (eigenvalues inside the bounds)."""
    emin, emax = float(bounds[0]), float(bounds[1])
    lam, vec = np.linalg.eigh(H)
    x = (2.0 * lam - emin - emax) / (emax - emin)
    if np.abs(x).max() > 1.0:
        raise ValueError("bounds do not contain the spectrum")
    w = (vec.T @ q0) ** 2
    th = np.arccos(np.clip(x, -1.0, 1.0))
    k = np.arange(nk)[:, None]
    return (np.cos(k * th[None, :]) * w[None, :]).sum(axis=1)


def fdstates(matvec, q0: np.ndarray, bounds, coef: np.ndarray):
    """(energies, vectors ndim x nsel) of the filter-diagonalisation states."""
    emin, emax = float(bounds[0]), float(bounds[1])
    de = emax - emin
    a, b = 2.0 / de, -(2.0 * emin / de + 1.0)
    coef = np.asarray(coef, dtype=np.float64)
    q0 = np.asarray(q0, dtype=np.float64)
    V = np.outer(q0, coef[0])
    qkm1, qkm2 = q0.copy(), None
    for k in range(1, coef.shape[0]):
        qk = a * matvec(qkm1) + b * qkm1
        if k > 1:
            qk = 2.0 * qk - qkm2
        V += np.outer(qk, coef[k])
        qkm2, qkm1 = qkm1, qk
    V /= np.sqrt((V * V).sum(axis=0))
    E = np.array([V[:, i] @ matvec(V[:, i]) for i in range(V.shape[1])])
    return E, V
