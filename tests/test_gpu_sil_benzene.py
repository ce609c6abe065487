"""SIL propagation on the benzene-size Hamiltonian (o=16, v=200, naux=864).

The autocorrelation of SIL is checked against an independent evaluation of
the same propagator, the Chebyshev expansion

    a(t) = F^T exp(-iHt) F = exp(-i b t) sum_k (2 - delta_k0) (-i)^k J_k(a t) C_k,
    C_k = F^T T_k((H - b)/a) F,  a = (Emax - Emin)/2,  b = (Emax + Emin)/2,

with the C_k from the GPU Chebyshev recursion (adc2g_cheby_auto) and the
Bessel functions J_k from scipy.  The SIL run of Lanczos order 12 must agree
with order 20, and the FP32-MFMA sigma-vector run with the FP64 one; the
differences are printed.  The initial state is the seeded synthetic dipole
vector (1h1p block, N(0, 0.1^2)), normalised; records every 0.1 a.u. to 2 a.u.
with the SIL tolerance 1e-10.
"""
import numpy as np
import pytest
from scipy.special import jv

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import SIZES, make_system, unit_dipole

pytestmark = pytest.mark.gpu

TOUT, TFINAL, TOL = 0.1, 2.0, 1e-10


def lanczos_extremes(op, steps=30, seed=1):
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(op.ndim)
    q /= np.linalg.norm(q)
    qp, b = np.zeros(op.ndim), 0.0
    alpha, beta = [], []
    for _ in range(steps):
        w = op.sigma(q) - b * qp
        a = q @ w
        w -= a * q
        b = np.linalg.norm(w)
        alpha.append(a)
        beta.append(b)
        qp, q = q, w / b
    ev = np.linalg.eigvalsh(np.diag(alpha) + np.diag(beta[:-1], 1) + np.diag(beta[:-1], -1))
    return float(ev[0]), float(ev[-1])


@pytest.fixture(scope="module")
def benzene_sil():
    sz = SIZES["benzene"]
    print(f"[case] SIL benzene {sz}", flush=True)
    sys = make_system(seed=0, **sz)
    op = lib.Operator.from_system(sys, keep_lowprec=True)
    f = np.zeros(op.ndim)
    f[: op.n1] = unit_dipole(sz["nocc"], sz["nvirt"], seed=0)
    f /= np.linalg.norm(f)
    yield op, f
    op.close()


def test_sil_benzene_vs_chebyshev(benzene_sil):
    op, f = benzene_sil
    print("[case] SIL benzene fp64 kdim 20", flush=True)
    a20, st = op.sil_autocorr(f, tout=TOUT, tfinal=TFINAL, kdim=20, autotol=TOL)
    nrec = int(TFINAL / (0.5 * TOUT))
    times = TOUT * np.arange(a20.size)
    assert a20.size == nrec + 1
    assert abs(a20[0] - 1.0) < 1e-14
    assert np.abs(a20).max() <= 1.0 + 1e-10
    # Chebyshev reference: bounds padded around 30-step Lanczos extremes
    lo, hi = lanczos_extremes(op)
    pad = 0.05 * (hi - lo)
    emin, emax = lo - pad, hi + pad
    ha, hb = 0.5 * (emax - emin), 0.5 * (emax + emin)
    kmax = int(ha * times[-1]) + 300
    order = (kmax + 1) // 2 + 1
    print(f"[case] Chebyshev reference: bounds ({emin:.2f}, {emax:.2f}) Eh, {2 * order + 1} moments", flush=True)
    ck = op.cheby_auto(f, (emin, emax), order)
    k = np.arange(ck.size)
    wk = np.where(k == 0, 1.0, 2.0) * (-1j) ** k
    ref = np.array([np.exp(-1j * hb * t) * np.sum(wk * jv(k, ha * t) * ck) for t in times])
    # the series is converged: the last Bessel weights are negligible
    assert np.abs(jv(k[-20:], ha * times[-1])).max() < 1e-14
    err = np.abs(a20 - ref).max()
    print(f"SIL order 20 vs Chebyshev: max |a_sil - a_cheb| = {err:.3e} ({st['nsteps']} SIL steps)")
    assert err < 1e-8


def test_sil_benzene_orders_and_precision(benzene_sil):
    op, f = benzene_sil
    a20, _ = op.sil_autocorr(f, tout=TOUT, tfinal=TFINAL, kdim=20, autotol=TOL)
    print("[case] SIL benzene fp64 kdim 12", flush=True)
    a12, st12 = op.sil_autocorr(f, tout=TOUT, tfinal=TFINAL, kdim=12, autotol=TOL)
    e12 = np.abs(a12 - a20).max()
    print(f"kdim 12 vs kdim 20: max |da| = {e12:.3e} ({st12['nsteps']} SIL steps)")
    assert e12 < 1e-8
    print("[case] SIL benzene fp32 kdim 20", flush=True)
    a32, _ = op.sil_autocorr(f, tout=TOUT, tfinal=TFINAL, kdim=20, autotol=TOL, precision="fp32")
    e32 = np.abs(a32 - a20).max()
    print(f"FP32-MFMA sigma vs FP64: max |da| = {e32:.3e}")
    assert e32 < 1e-6
