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
* ``propagate`` follows propagate_wavepacket_sil,
  source/propagation/proplib_adc2.f90:486-630: periods of length tout, SIL
  steps (oracle/sil.py silstep) with H at each step's start time, a period
  skipped when |dpsi/dt| <= 1e-10, norms at the start of each period and at
  the end.

Pinned by (tests/test_oracle_pins.py): the envelopes' half-maximum widths
(the fwhm parameter of the cos^2 and sin^2 envelopes), ramp and box values;
a constant field (box envelope, zero frequency), where the propagation must
equal exp(-i (H - E.D) t) psi0 of the (ndim + 1)-dimensional matrix from a
dense exponential; norm conservation under a time-dependent field.
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


def matxvec_treal_laser(hmul, dips, E, v):
    """-i H(t) v.  hmul(x) = H x for complex x of ndim components; dips[c] =
    (D_c dense (ndim x ndim), d00_c, d0j_c) or None."""
    nd = v.size - 1
    out = np.zeros_like(v)
    out[:nd] = -1j * hmul(v[:nd])
    for c in range(3):
        if E[c] == 0.0:
            continue
        D, d00, d0j = dips[c]
        out[:nd] += 1j * E[c] * (D @ v[:nd] + d00 * v[:nd] + d0j * v[nd])
        out[nd] += 1j * E[c] * (d00 * v[nd] + np.dot(d0j, v[:nd]))
    return out


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


def propagate(hmul, dips, pulses, psi0, tout, tfinal, kdim, proptol):
    """Returns (psi(tfinal), norms at i * tout, i = 0 .. int(tfinal / tout))."""
    tiny, tinier = 1e-9, 1e-10
    psi = np.array(psi0, dtype=np.complex128)
    intperiod = tout
    nrec = int(tfinal / intperiod)
    norms = np.zeros(nrec + 1)
    time = 0.0
    for i in range(1, nrec + 1):
        inttime = 0.0
        while True:
            step = intperiod - inttime
            t_now = time
            func = (lambda x, t_now=t_now: matxvec_treal_laser(hmul, dips, efield(pulses, t_now), x))
            dtpsi = func(psi)
            if inttime == 0.0:
                norms[i - 1] = np.sqrt(np.vdot(psi, psi).real)
            if np.sqrt(np.vdot(dtpsi, dtpsi).real) > tinier:
                psi, truestep, _ = sil.silstep(psi, dtpsi, func, step, kdim, proptol)
                time += truestep
                inttime += truestep
                if abs(intperiod - inttime) > abs(tiny * intperiod):
                    continue
            else:
                time += intperiod
            break
    norms[nrec] = np.sqrt(np.vdot(psi, psi).real)
    return psi, norms
