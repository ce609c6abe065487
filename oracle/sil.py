"""Short-iterative-Lanczos propagation restated.  Test oracle only.

* ``silstep`` follows ``SILStep`` of source/propagation/sillib.f90:108-393 for
  the path the ADC driver uses (Restart = .true., StdForm = .true.,
  Relax = .false., so PreFactor = -i): first Lanczos element and error
  estimate |beta*dt| (:198-230), the Krylov loop with
  Err_k = |Err_{k-1} * beta_k * dt / k| (:234-323), the next diagonal element
  estimated as the previous one (:327), stepsize reduction
  dt * (TolError/Err)^(1/order) (:331-334), DSTEQR on the order+1 tridiagonal
  (:350-361) and the exponential combination (:369-390).  Error codes 1-4 as
  in :84-89.
* ``propagate`` follows ``propagate_sillib`` of
  source/propagation/fvecprop.f90:365-564: psi0 = F/|F| (init_wavepacket
  :338-357), intperiod = tout/2, records i = 1..int(tfinal/intperiod) with the
  t/2 trick a0(2 t) = sum_D psi_D(t)^2 * |F|^2 (:521) at time i*tout.

* ``fdstates_td`` follows calc_fdstates_tdfd of
  source/propagation/fdstates_tdfd.f90:28-627 (periods of tout, the windowed
  filter states and their combination).

``func(v)`` must return -i H v for complex v (``matxvec_treal``,
source/propagation/tdselib.f90:323-437).

Pinned by: the exact propagator exp(-iHt) from a dense eigendecomposition at
small sizes, norm conservation and the t/2 identity (test_oracle_pins); the
filter-diagonalisation states by the same states built from the exact
psi(t) (``states`` argument), and by a filter basis concentrated on one
isolated eigenvalue, which returns that eigenvalue.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


class SILError(RuntimeError):
    def __init__(self, code):
        msgs = {1: "Illegal integration order", 2: "Stepsize underflow",
                3: "Diagonalisation failed", 4: "Illegal initial stepsize"}
        super().__init__(msgs.get(code, "Unknown error occurred"))
        self.code = code


def tridiag_eig(diag, off):
    """Eigenpairs of the symmetric tridiagonal (diag, off); DSTEQR('I')."""
    w, U = scipy.linalg.eigh_tridiagonal(diag, off)
    return w, U


def silstep(psi, dtpsi, func, intperiod, intorder, tolerror):
    """One SILStep; returns (psi_new, stepsize, trueorder).  ``dtpsi`` is
    -i H psi on entry (consumed)."""
    RelativeMinStep = 1.0e-10
    if intorder < 2:
        raise SILError(1)
    stepsize = intperiod
    minstep = abs(RelativeMinStep * intperiod)
    pref = -1j
    psi = psi.copy()
    dtpsi = dtpsi.copy()
    diag = np.zeros(intorder + 1)
    offd = np.zeros(intorder + 1)
    kry = [psi, None]
    alpha = np.vdot(psi, dtpsi)
    diag[0] = (alpha / pref).real
    dtpsi = dtpsi - alpha * psi
    beta = np.sqrt(np.vdot(dtpsi, dtpsi).real)
    offd[1] = beta
    cbeta = pref * beta
    if beta < 1.0e-20:
        beta = 0.0
        cinv = 0.0
    else:
        cinv = 1.0 / cbeta
    dtpsi = cinv * dtpsi
    kry[1] = dtpsi
    error = abs(beta * stepsize)
    trueorder = 1
    while True:
        if abs(stepsize) < abs(minstep):
            raise SILError(2)
        trueorder += 1
        w = func(kry[trueorder - 1])
        alpha = np.vdot(kry[trueorder - 1], w)
        diag[trueorder - 1] = (alpha / pref).real
        w = w - alpha * kry[trueorder - 1] - cbeta * kry[trueorder - 2]
        beta = np.sqrt(np.vdot(w, w).real)
        offd[trueorder] = beta
        cbeta = pref * beta
        if beta < 1.0e-20:
            offd[trueorder] = 0.0
            beta = 0.0
            cinv = 0.0
        else:
            cinv = 1.0 / cbeta
        kry.append(cinv * w)
        error = abs(error * beta * stepsize / float(trueorder))
        if not (trueorder < intorder and error > tolerror and beta > 1.0e-20):
            break
    diag[trueorder] = diag[trueorder - 1]
    if error > tolerror:
        stepsize = stepsize * (tolerror / error) ** (1.0 / float(trueorder))
    lam, U = tridiag_eig(diag[:trueorder + 1], offd[1:trueorder + 1])
    coef = (np.exp(pref * lam * stepsize)[None, :] * U[0][None, :] * U).sum(axis=1)
    out = np.zeros_like(psi)
    for p in range(trueorder + 1):
        out = out + coef[p] * kry[p]
    return out, stepsize, trueorder


def propagate(func, fvec, tout, tfinal, kdim, autotol, record_orders=False, autoord=0):
    """Return (times, auto0) with auto0[0] at t=0 and auto0[i] at i*tout.

    With autoord = 1 or 2, auto0 is replaced by the list [auto0, auto1, ...]
    of propagate_sillib (source/propagation/fvecprop.f90:421-541):
    auto_k[0] = dot_product(psi0, H^k psi0) |F|^2 (conjugating psi0) and
    auto_k[i] = sum(psi H^k psi) |F|^2 (no conjugation: the t/2 trick), with
    H v = i func(v)."""
    norm0 = np.sqrt(np.vdot(fvec, fvec).real)
    psi = fvec.astype(np.complex128) / norm0
    intperiod = 0.5 * tout
    tiny = 1e-9
    nrec = int(tfinal / intperiod)
    auto = np.empty(nrec + 1, dtype=np.complex128)
    auto[0] = np.vdot(psi, psi) * norm0 ** 2
    higher = [np.empty(nrec + 1, dtype=np.complex128) for _ in range(autoord)]
    hk = psi
    for k in range(autoord):
        hk = 1j * func(hk)
        higher[k][0] = np.vdot(psi, hk) * norm0 ** 2
    orders = []
    for i in range(1, nrec + 1):
        inttime = 0.0
        while True:
            stepsize = intperiod - inttime
            dtpsi = func(psi)
            psi, truestep, order = silstep(psi, dtpsi, func, stepsize, kdim, autotol)
            orders.append(order)
            inttime += truestep
            if not (abs(intperiod - inttime) > abs(tiny * intperiod)):
                break
        auto[i] = np.sum(psi * psi) * norm0 ** 2
        hk = psi
        for k in range(autoord):
            hk = 1j * func(hk)
            higher[k][i] = np.sum(psi * hk) * norm0 ** 2
    times = np.arange(nrec + 1) * tout
    if autoord:
        auto = [auto] + higher
    if record_orders:
        return times, auto, orders
    return times, auto


def fdstates_td(func, hmul, fvec, tout, tfinal, kdim, autotol, ebound, f2e, iwfunc, states=None):
    """calc_fdstates_tdfd (source/propagation/fdstates_tdfd.f90:28-627):
    psi0 = fvec / |fvec|, SIL periods of tout (calc_filterstates,
    :315-461), the filter states buffer_j += fac dt g(t) Re(exp(i E_j t)
    psi(t)) at every record (update_filterstates, :465-520; windowfunc
    cos(pi t / (2 tfinal))^iwfunc, :524-535), v_i = sum_j f2e[i, j] buffer_j,
    normalised, E_i = v_i . H v_i (:556-627).  func(x) = -i H x, hmul(x) =
    H x.  With ``states`` (a function t -> psi(t)) the wavefunctions come
    from it instead of the SIL steps.  Returns (energies, vectors)."""
    f2e = np.atleast_2d(np.asarray(f2e, dtype=np.float64))
    nsel, nfbas = f2e.shape
    dt = tout
    nrec = int(tfinal / dt)
    de = (ebound[1] - ebound[0]) / (nfbas - 1)
    E = ebound[0] + de * np.arange(nfbas)
    psi = fvec.astype(np.complex128) / np.sqrt(np.vdot(fvec, fvec).real)
    buf = np.zeros((fvec.size, nfbas))
    tiny = 1e-9

    def update(psi, r):
        t = r * dt
        fac = 1.0 if (r == 0 or r == nrec) else 2.0
        gk = np.cos(np.pi * t / (2.0 * tfinal)) ** iwfunc
        buf[:] += fac * dt * gk * np.real(np.exp(1j * E[None, :] * t) * psi[:, None])

    update(psi if states is None else states(0.0), 0)
    for i in range(1, nrec + 1):
        if states is None:
            inttime = 0.0
            while True:
                stepsize = dt - inttime
                psi, truestep, _ = silstep(psi, func(psi), func, stepsize, kdim, autotol)
                inttime += truestep
                if not (abs(dt - inttime) > abs(tiny * dt)):
                    break
            update(psi, i)
        else:
            update(states(i * dt), i)
    V = buf @ f2e.T
    V /= np.sqrt((V * V).sum(axis=0))
    Es = np.array([V[:, k] @ hmul(V[:, k]) for k in range(nsel)])
    return Es, V
