"""GPU SIL propagation against the oracle restatement and the exact propagator.

The GPU follows the same SILStep decisions as the oracle (oracle/sil.py), so
the autocorrelation functions agree to 1e-10 relative to |F|^2; against the
exact exp(-iHt) they agree to the SIL tolerance.  The FP32-MFMA sigma variant
is compared with the FP64 run (error reported; bar 1e-5 relative).
"""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system, unit_dipole
from oracle import dfsigma, sil

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    print("[case] sil small o=3 v=6 naux=8", flush=True)
    sys = make_system(3, 6, 8, seed=5)
    df = dfsigma.DF(sys)
    H = df.sigma(np.eye(df.n1 + df.D2.size))
    op = lib.Operator.from_system(sys)
    f = np.zeros(op.ndim)
    f[: op.n1] = unit_dipole(3, 6, seed=1)
    yield sys, H, op, f
    op.close()


def test_matxvec_treal(small):
    sys, H, op, f = small
    rng = np.random.default_rng(3)
    v = rng.standard_normal(op.ndim) + 1j * rng.standard_normal(op.ndim)
    out = op.matxvec_treal(v)
    ref = -1j * (H @ v)
    assert np.abs(out - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("kdim,tol", [(12, 1e-12), (6, 1e-8), (20, 1e-10)])
def test_sil_vs_oracle_and_exact(small, kdim, tol):
    sys, H, op, f = small
    auto, st = op.sil_autocorr(f, tout=0.1, tfinal=3.0, kdim=kdim, autotol=tol)
    times, ref = sil.propagate(lambda x: -1j * (H @ x), f, tout=0.1, tfinal=3.0, kdim=kdim, autotol=tol)
    nf = np.dot(f, f)
    assert auto.size == ref.size
    assert np.abs(auto - ref).max() < 1e-10 * nf
    w, U = np.linalg.eigh(H)
    c = U.T @ f
    exact = np.array([np.sum(c * c * np.exp(-1j * w * t)) for t in times])
    assert np.abs(auto - exact).max() < max(100 * tol, 1e-10) * nf * 10
    assert st["max_order"] <= kdim


def test_sil_water_vs_oracle():
    print("[case] sil water-size o=5 v=36 naux=164", flush=True)
    sys = make_system(5, 36, 164, seed=0)
    df = dfsigma.DF(sys)
    with lib.Operator.from_system(sys) as op:
        f = np.zeros(op.ndim)
        f[: op.n1] = unit_dipole(5, 36, seed=0)
        auto, st = op.sil_autocorr(f, tout=0.1, tfinal=2.0, kdim=20, autotol=1e-10)
        _, ref = sil.propagate(lambda x: -1j * df.sigma(x.real) + df.sigma(x.imag), f, tout=0.1, tfinal=2.0,
                               kdim=20, autotol=1e-10)
        nf = np.dot(f, f)
        err = np.abs(auto - ref).max() / nf
        print("max |a_gpu - a_oracle| / |F|^2 =", err)
        assert err < 1e-8
        a32, _ = op.sil_autocorr(f, tout=0.1, tfinal=2.0, kdim=20, autotol=1e-10, precision="fp32")
        e32 = np.abs(a32 - auto).max() / nf
        print("fp32-MFMA sigma: max |a_fp32 - a_fp64| / |F|^2 =", e32)
        assert e32 < 1e-5


def test_sil_bad_order(small):
    sys, H, op, f = small
    with pytest.raises(lib.Adc2gError) as e:
        op.sil_autocorr(f, tout=0.1, tfinal=0.2, kdim=1, autotol=1e-8)
    assert e.value.code == 1


@pytest.mark.parametrize("autoord", [1, 2])
def test_sil_higher_orders_vs_oracle_and_exact(small, autoord):
    """Orders 1 and 2 of propagate_sillib against the oracle (1e-10 relative
    to sum c_n^2 |E_n|^k) and the closed form sum_n c_n^2 E_n^k exp(-i E_n t)."""
    sys, H, op, f = small
    autos, st = op.sil_autocorr_orders(f, tout=0.1, tfinal=3.0, kdim=12, autotol=1e-12, autoord=autoord)
    times, ref = sil.propagate(lambda x: -1j * (H @ x), f, tout=0.1, tfinal=3.0, kdim=12, autotol=1e-12,
                               autoord=autoord)
    w, U = np.linalg.eigh(H)
    c = U.T @ f
    assert len(autos) == autoord + 1
    for k in range(autoord + 1):
        scale = np.sum(c * c * np.abs(w) ** k)
        assert np.abs(autos[k] - ref[k]).max() < 1e-10 * scale, k
        exact = np.array([np.sum(c * c * w ** k * np.exp(-1j * w * t)) for t in times])
        assert np.abs(autos[k] - exact).max() < 1e-8 * scale, k
    # the zeroth order equals the plain entry point's
    auto0, _ = op.sil_autocorr(f, tout=0.1, tfinal=3.0, kdim=12, autotol=1e-12)
    assert np.abs(autos[0] - auto0).max() == 0.0


@pytest.mark.parametrize("tfinal,iwfunc", [(3.0, 1), (20.0, 2)])
def test_sil_fdstates_vs_oracle(tfinal, iwfunc):
    """Time-domain filter-diagonalisation states (calc_fdstates_tdfd):
    energies and states against the oracle's SIL propagation (pinned by the
    exact propagator), with fewer records than the 16-record ring and with
    records across it."""
    from oracle import dfsigma
    print(f"[case] sil fdstates o=3 v=5 tfinal={tfinal} iwfunc={iwfunc}", flush=True)
    sys_ = make_system(3, 5, 8, seed=6)
    df = dfsigma.DF(sys_)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    lam = np.linalg.eigvalsh(H)
    rng = np.random.default_rng(6)
    f = np.zeros(nd)
    f[: df.n1] = rng.normal(size=df.n1)
    F = rng.normal(size=(3, 9))
    eb = (lam[0], lam[0] + 0.5 * (lam[-1] - lam[0]))
    E_ref, V_ref = sil.fdstates_td(lambda x: -1j * (H @ x), lambda x: H @ x, f, 0.5, tfinal, 14, 1e-10, eb, F,
                                   iwfunc)
    with lib.Operator.from_system(sys_) as op:
        E, V, st = op.sil_fdstates(f, 0.5, tfinal, 14, 1e-10, eb, F, iwfunc=iwfunc)
    print(f"{st['nsteps']} SIL steps, energies {np.round(E, 6)}", flush=True)
    assert np.abs(V - V_ref).max() < 1e-8
    assert np.abs(E - E_ref).max() < 1e-8
