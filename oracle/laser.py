"""Laser-driven SIL propagation restated.  Test oracle only.

* ``efield`` follows ``efield`` and its envelopes,
  source/propagation/tdselib.f90:850-1029: E(t) = sum over pulses of
  pulse_vec * strength * envelope(t) * carrier(t), carrier 1 cos(w (t - t0) -
  phase) or 2 sin of the same; envelopes 1 cos^2 (:904-929), 2 sin^2 ramp
  (:968-993), 3 box (:999-1024), 4 sin^2 (:936-961).
* ``matxvec_treal_laser`` follows tdselib.f90:569-720 without a CAP: the
  wavefunction has ndim + 1 components (ground state last); -i H on the
  intermediate states (matxvec_treal), plus i E_c [(D_c + d00_c) v + d0j_c v0]
  on them and i E_c (d00_c v0 + d0j_c . v) on the ground state.
* the CAP term of tdselib.f90:722-830: -P [(W + w00) u + w0j u0] on the
  intermediate states and -P (w00 u0 + w0j . u) on the ground state, u = P v,
  P the projector onto the complement of the selected bound states (vectors
  and/or the ground state).
* ``csilstep`` follows CSILStep, source/propagation/csillib.f90:120-520, with
  the improved error criterion (OldErrCri = .false.), Restart = .true.,
  StdForm = .true.: Arnoldi with modified Gram-Schmidt, Hessenberg of -i H,
  the last column estimated (sub-diagonal mirrored, diagonal repeated), the
  coefficients exp(Hs dt) e_0 through the eigen-decomposition of Hs / (-i)
  and a linear solve (eigchess, lineqz), the error the change of the
  coefficients between orders, the step reduced by (tol / former error)^(1/k).
* ``flux`` follows adc2_flux_cap, source/propagation/flux.f90:179-411:
  2 Re <dpsi/dt| P Theta' P |psi> + 2 Re <psi| P W' P |psi>, where Theta'
  (W') on the intermediate states is (1 + theta00) Theta_IS u + theta0j u0 as
  the reference writes it (oppsi = oppsi + theta00 * oppsi, :263 and :354)
  and theta00 u0 + theta0j . u on the ground state.
* ``propagate`` follows propagate_wavepacket_sil,
  source/propagation/proplib_adc2.f90:486-630 (SIL steps, oracle/sil.py
  silstep), or with a CAP propagate_wavepacket_csil, :659-830 (csilstep):
  periods of length tout, steps with H at each step's start time, a period
  skipped when |dpsi/dt| <= 1e-10, norms at the start of each period and at
  the end; with Theta the flux at the same times (:766, :816-819), the last
  one from dpsi/dt at the final time.

Pinned by (tests/test_oracle_pins.py): the envelopes' half-maximum widths
(the fwhm parameter of the cos^2 and sin^2 envelopes), ramp and box values;
a constant field (box envelope, zero frequency), where the propagation must
equal exp(-i (H - E.D) t) psi0 of the (ndim + 1)-dimensional matrix from a
dense exponential; norm conservation under a time-dependent field; with a
CAP the constant-field propagation equals exp(-i (H - E.D - i P W P) t) psi0
and the norm decays; the flux with theta00 = w00 = 0 equals
d/dt [<psi|P Theta P|psi> - <psi|psi>] by finite differences of that exact
propagation, and with Theta = 2 (W = c) on the intermediate states and no
off-diagonal terms it reduces to 2 (1 + theta00) d/dt |psi_IS|^2 +
theta00 d/dt |psi_0|^2 (2 (1 + w00) c |psi_IS|^2 + 2 w00 |psi_0|^2), which
fixes where theta00 and w00 enter.
"""
from __future__ import annotations

import numpy as np

from . import sil


def envelope(p: dict, t: float) -> float:
    kind, ep = p["envelope"], list(p.get("envpar", (0.0, 0.0))) + [0.0, 0.0]
    if kind == "cos2":
        tz, fw = ep[0], ep[1]
        return float(np.cos(np.pi * (t - tz) / (2.0 * fw)) ** 2) if tz - fw < t < tz + fw else 0.0
    if kind == "sin2ramp":
        tau = ep[0]
        if t < 0.0:
            return 0.0
        return float(np.sin(np.pi * t / (2.0 * tau)) ** 2) if t <= tau else 1.0
    if kind == "box":
        return 0.0 if (t < ep[0] or t > ep[1]) else 1.0
    if kind == "sin2":
        tz, fw = ep[0], ep[1]
        return float(np.sin(np.pi * (t - tz) / (2.0 * fw)) ** 2) if tz < t < tz + 2.0 * fw else 0.0
    raise ValueError(kind)


def efield(pulses, t: float) -> np.ndarray:
    E = np.zeros(3)
    for p in pulses:
        arg = p.get("freq", 0.0) * (t - p.get("t0", 0.0)) - p.get("phase", 0.0)
        carrier = np.cos(arg) if p.get("carrier", "cos") == "cos" else np.sin(arg)
        E += np.asarray(p["pulse_vec"], dtype=float) * p["strength"] * envelope(p, t) * carrier
    return E


def dense_dipole(diag, row, col, val, ndim):
    """The symmetric IS matrix of the reference's operator file layout."""
    D = np.diag(np.asarray(diag, dtype=float))
    np.add.at(D, (row, col), val)
    np.add.at(D, (col, row), val)
    return D


def _project(x, proj):
    """P x: proj = (R (ndim x nvec) orthonormal or None, ground flag)."""
    R, ground = proj
    x = x.copy()
    nd = x.size - 1
    if R is not None and R.shape[1] > 0:
        x[:nd] -= R @ (R.T @ x[:nd])
    if ground:
        x[nd] = 0.0
    return x


def matxvec_treal_laser(hmul, dips, E, v, cap=None, proj=(None, False)):
    """-i H(t) v.  hmul(x) = H x for complex x of ndim components; dips[c] =
    (D_c dense (ndim x ndim), d00_c, d0j_c) or None; cap = (W dense, w00,
    w0j) or None; proj = (R, ground) the CAP projector."""
    nd = v.size - 1
    out = np.zeros_like(v)
    out[:nd] = -1j * hmul(v[:nd])
    for c in range(3):
        if E[c] == 0.0:
            continue
        D, d00, d0j = dips[c]
        out[:nd] += 1j * E[c] * (D @ v[:nd] + d00 * v[:nd] + d0j * v[nd])
        out[nd] += 1j * E[c] * (d00 * v[nd] + np.dot(d0j, v[:nd]))
    if cap is not None:
        W, w00, w0j = cap
        u = _project(v, proj)
        t = np.zeros_like(v)
        t[:nd] = -(W @ u[:nd]) - w00 * u[:nd] - w0j * u[nd]
        t[nd] = -w00 * u[nd] - np.dot(w0j, u[:nd])
        out += _project(t, proj)
    return out


def cap_matrix(cap, proj, nd):
    """P W_full P on the ndim + 1 states (W_full = [[W + w00, w0j], [w0j^T, w00]])."""
    W, w00, w0j = cap
    Wf = np.zeros((nd + 1, nd + 1))
    Wf[:nd, :nd] = W + w00 * np.eye(nd)
    Wf[:nd, nd] = w0j
    Wf[nd, :nd] = w0j
    Wf[nd, nd] = w00
    P = np.eye(nd + 1)
    R, ground = proj
    if R is not None and R.shape[1] > 0:
        P[:nd, :nd] -= R @ R.T
    if ground:
        P[nd, nd] = 0.0
    return P @ Wf @ P


def _flux_op(op, proj, x):
    """P Op' P x for op = (M dense IS block, m00, m0j) as adc2_flux_cap
    applies it (flux.f90:257-272)."""
    M, m00, m0j = op
    nd = x.size - 1
    u = _project(x, proj)
    o = np.zeros_like(u)
    o[:nd] = M @ u[:nd]
    o[:nd] = o[:nd] + m00 * o[:nd]
    o[nd] = m00 * u[nd] + np.dot(m0j, u[:nd])
    o[:nd] += m0j * u[nd]
    return _project(o, proj)


def flux(theta, cap, proj, psi, dtpsi):
    """The CAP flux; theta and cap as (M, m00, m0j)."""
    val1 = 2.0 * np.vdot(dtpsi, _flux_op(theta, proj, psi)).real
    val2 = 2.0 * np.vdot(psi, _flux_op(cap, proj, psi)).real
    return float(val1 + val2)


def csilstep(psi, dtpsi, func, intperiod, intorder, tolerror):
    """One CSILStep; returns (psi_new, stepsize, trueorder)."""
    RelativeMinStep, Tiny, MinOrder = 1.0e-9, 1.0e-18, 2
    if intorder < MinOrder or intorder > 50:
        raise sil.SILError(1)
    pref = -1j
    stepsize = intperiod
    minstep = abs(RelativeMinStep * intperiod)
    norm = np.sqrt(np.vdot(psi, psi).real)
    kry = [psi / norm, dtpsi / norm]
    Hs = np.zeros((intorder + 1, intorder + 1), dtype=complex)
    former, order = 1.0, 0
    old = np.zeros(intorder + 1, dtype=complex)

    def coeffs(m, dt):
        lam, V = np.linalg.eig(Hs[:m, :m] / pref)
        a2 = np.linalg.solve(V, np.eye(m)[:, 0])
        return V @ (np.exp(pref * lam * dt) * a2)

    while True:
        if abs(stepsize) < abs(minstep):
            raise sil.SILError(2)
        nxt = order + 1
        w = kry[nxt]
        for p in range(order + 1):
            c = np.vdot(kry[p], w)
            Hs[p, order] = c
            w = w - c * kry[p]
        sub = np.sqrt(np.vdot(w, w).real)
        Hs[nxt, order] = pref * sub
        if sub < Tiny:
            sub = 0.0
            w = 0.0 * w
        else:
            w = w / Hs[nxt, order]
        kry[nxt] = w
        order = nxt
        former = abs(former * stepsize * sub / order)
        Hs[order - 1, order] = Hs[order, order - 1]
        Hs[order, order] = Hs[order - 1, order - 1]
        c1 = coeffs(order + 1, stepsize)
        error = 0.0 if order == 1 else np.sqrt(np.sum(np.abs(old[: order + 1] - c1) ** 2))
        old[: order + 1] = c1
        if (error > tolerror and order < intorder) or order < MinOrder:
            kry.append(func(kry[order]))
            continue
        break
    if not error <= tolerror:
        stepsize = min(stepsize, stepsize * (tolerror / former) ** (1.0 / order))
        c1 = coeffs(order + 1, stepsize)
    new = sum(c1[p] * kry[p] for p in range(order + 1))
    return norm * new, stepsize, order


def full_matrix(H, dips, E):
    """H - E.D on the ndim + 1 states (the generator of the constant-field
    propagation): [[H - E.(D + d00), -E.d0j], [-E.d0j^T, -E.d00]]."""
    nd = H.shape[0]
    M = np.zeros((nd + 1, nd + 1))
    M[:nd, :nd] = H
    for c in range(3):
        if E[c] == 0.0:
            continue
        D, d00, d0j = dips[c]
        M[:nd, :nd] -= E[c] * (D + d00 * np.eye(nd))
        M[:nd, nd] -= E[c] * d0j
        M[nd, :nd] -= E[c] * d0j
        M[nd, nd] -= E[c] * d00
    return M


def propagate(hmul, dips, pulses, psi0, tout, tfinal, kdim, proptol, cap=None, proj=(None, False), theta=None):
    """Returns (psi(tfinal), norms at i * tout, i = 0 .. int(tfinal / tout)),
    and the flux at the same times after the norms when theta is given."""
    stepper = sil.silstep if cap is None else csilstep
    tiny, tinier = 1e-9, 1e-10
    psi = np.array(psi0, dtype=np.complex128)
    intperiod = tout
    nrec = int(tfinal / intperiod)
    norms = np.zeros(nrec + 1)
    fl = np.zeros(nrec + 1)
    time = 0.0
    for i in range(1, nrec + 1):
        inttime = 0.0
        while True:
            step = intperiod - inttime
            t_now = time
            func = (lambda x, t_now=t_now: matxvec_treal_laser(hmul, dips, efield(pulses, t_now), x, cap, proj))
            dtpsi = func(psi)
            if inttime == 0.0:
                norms[i - 1] = np.sqrt(np.vdot(psi, psi).real)
                if theta is not None:
                    fl[i - 1] = flux(theta, cap, proj, psi, dtpsi)
            if np.sqrt(np.vdot(dtpsi, dtpsi).real) > tinier:
                psi, truestep, _ = stepper(psi, dtpsi, func, step, kdim, proptol)
                time += truestep
                inttime += truestep
                if abs(intperiod - inttime) > abs(tiny * intperiod):
                    continue
            else:
                time += intperiod
            break
    norms[nrec] = np.sqrt(np.vdot(psi, psi).real)
    if theta is not None:
        dtpsi = matxvec_treal_laser(hmul, dips, efield(pulses, time), psi, cap, proj)
        fl[nrec] = flux(theta, cap, proj, psi, dtpsi)
        return psi, norms, fl
    return psi, norms
