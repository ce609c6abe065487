"""Ground-to-excited transition density (adc2g_trdens_vo), the natural
transition orbitals built from it (lib.nto) and the TD-NTO records of
adc2g_sil_laser, against the oracle (oracle/nto.py: adc2_trdens_gs_virt_occ
and tdadc2_nto, pinned in tests/test_oracle_pins.py as the transpose of the
F-vector's dpl_vo terms).  FP64 throughout: the density to 1e-12 relative,
the records of a propagation to 1e-8 (the propagated wavefunction agrees to
that order)."""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import dfsigma, laser, nto

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("o,v,naux,seed,form", [(2, 4, 6, 1, "df"), (3, 5, 8, 2, "int"), (4, 12, 20, 6, "df")])
def test_trdens_vo_vs_oracle(o, v, naux, seed, form):
    print(f"[case] trdens_vo o={o} v={v} naux={naux} coupling={form}", flush=True)
    sys_ = make_system(o, v, naux, seed=seed)
    rng = np.random.default_rng(seed)
    with lib.Operator.from_system(sys_, coupling=form) as op:
        X = rng.normal(size=(op.ndim, 3))
        got = op.trdens_vo(X)
        one = op.trdens_vo(X[:, 1])
        X2 = X.copy()
        X2[o * v:] = rng.normal(size=X2[o * v:].shape)    # the 2h2p rows do not enter
        again = op.trdens_vo(X2)
    ref = np.stack([nto.trdens_vo(sys_, X[:, s]) for s in range(3)])
    assert got.shape == (3, v, o)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.array_equal(one, got[1])
    assert np.array_equal(again, got)


def test_trdens_vo_cvs_view_and_bad_sizes():
    """On a CVS view the view's 1h1p amplitudes (core holes) enter at their
    configurations with the full MP2 density (rho_mp2 changes only with
    frozen core)."""
    print("[case] trdens_vo on a CVS view", flush=True)
    sys_ = make_system(3, 5, 8, seed=3)
    o, v = sys_.nocc, sys_.nvirt
    rng = np.random.default_rng(3)
    with lib.Operator.from_system(sys_) as op:
        with op.cvs([1, 3]) as cv:
            Xv = rng.normal(size=(cv.ndim, 2))
            T = cv.trdens_vo(Xv)
            n1v = cv.n1
        with pytest.raises(ValueError):
            op.trdens_vo(np.ones(op.ndim - 1))
    rows = np.array([a * o + i for a in range(v) for i in (0, 2)])
    assert rows.size == n1v
    for s in range(2):
        x = np.zeros(o * v)
        x[rows] = Xv[:n1v, s]
        ref = nto.trdens_vo(sys_, x)
        assert np.abs(T[s] - ref).max() < 1e-12 * np.abs(ref).max(), s


def test_nto_of_the_gpu_density():
    """lib.nto of the GPU density: singular values, pair count and the
    density rebuilt from the pairs against tdadc2_nto of the oracle."""
    print("[case] nto o=3 v=5", flush=True)
    sys_ = make_system(3, 5, 8, seed=4)
    o, v = sys_.nocc, sys_.nvirt
    rng = np.random.default_rng(4)
    with lib.Operator.from_system(sys_) as op:
        wf = rng.normal(size=op.ndim) + 1j * rng.normal(size=op.ndim)
        wf[0] *= 8.0                                      # one dominant pair
        w1 = wf / np.linalg.norm(wf)
        T = op.trdens_vo(w1.real) + 1j * op.trdens_vo(w1.imag)
    sigma, hole, particle = lib.nto(T)
    Tr, sr, Ur, Vr, npair = nto.tdnto(sys_, wf)
    assert np.abs(T - Tr).max() < 1e-12 * np.abs(Tr).max()
    assert np.abs(sigma - sr).max() < 1e-12 * sr.max()
    assert hole.shape == (o, npair) and particle.shape == (v, npair)
    full_s, full_h, full_p = lib.nto(T, thresh=0.0)
    rebuilt = (full_h * full_s) @ full_p.conj().T
    assert np.abs(rebuilt - T.T).max() < 1e-12 * np.abs(T).max()
    C = rng.normal(size=(7, o))
    assert np.allclose(lib.nto(T, c_occ=C)[1], C @ hole)


def test_sil_laser_trdens_records():
    """The transition density of psi_IS / |psi_IS| at every tout of a laser
    propagation equals tdadc2_nto's density of the oracle's wavefunction at
    the same time."""
    print("[case] sil laser trdens records o=3 v=5", flush=True)
    sys_ = make_system(3, 5, 8, seed=5)
    df = dfsigma.DF(sys_)
    nd = df.n1 + df.D2.size
    H = df.sigma(np.eye(nd))
    H = 0.5 * (H + H.T)
    rng = np.random.default_rng(5)
    row = rng.integers(0, nd, 4 * nd)
    col = (row + 1 + rng.integers(0, nd - 1, row.size)) % nd
    val = rng.normal(0.0, 0.05, row.size)
    diag = rng.normal(0.0, 0.2, nd)
    d0j = rng.normal(0.0, 0.1, nd)
    dips = [(laser.dense_dipole(diag, row, col, val, nd), 0.2, d0j), None, None]
    pulses = [{"envelope": "cos2", "envpar": (1.0, 1.5), "carrier": "cos", "freq": 0.8,
               "pulse_vec": (1.0, 0.0, 0.0), "strength": 0.2}]
    psi0 = np.zeros(nd + 1, dtype=complex)
    psi0[nd] = 0.9
    psi0[:nd] = 0.05 * rng.normal(size=nd)
    with lib.Operator.from_system(sys_) as op:
        op.set_dipole_is(0, diag, row, col, val, 0.2, d0j)
        psi, norms, T, st = op.sil_laser(psi0, pulses, tout=0.5, tfinal=2.0, kdim=16, proptol=1e-10, trdens=True)
    assert T.shape == (5, sys_.nvirt, sys_.nocc)
    for i in range(5):
        ref_psi = laser.propagate(lambda z: H @ z, dips, pulses, psi0, 0.5, 0.5 * i, 16, 1e-10)[0] if i else psi0
        Tr = nto.tdnto(sys_, ref_psi[:nd])[0]
        assert np.abs(T[i] - Tr).max() < 1e-8 * np.abs(Tr).max(), i
    assert np.abs(T[-1] - T[0]).max() > 1e-3                # the pulse changed the density


def test_frozen_core_mp2_and_trdens():
    """Frozen-core view (lifrzcore): adc2g_mp2 equals the MP2 of the system
    without the frozen orbitals (energy, D2) and its density embedded, which
    is the oracle's frozen-core density; the transition density of view
    vectors equals the oracle's with the view's 1h1p amplitudes at their
    configurations and that density."""
    from oracle import mp2
    print("[case] frozen-core mp2 and trdens o=5 v=9 frozen 1,3", flush=True)
    sys_ = make_system(5, 9, 12, seed=8)
    o, v = sys_.nocc, sys_.nvirt
    frozen = [1, 3]                                    # 1-based occupied orbitals
    keep = [i for i in range(o) if i + 1 not in frozen] + list(range(o, o + v))
    rng = np.random.default_rng(8)
    with lib.Operator.from_system(sys_) as op, op.frozen_core(frozen) as fc:
        e, d2, rho = fc.mp2()
        Xv = rng.normal(size=(fc.ndim, 2))
        T = fc.trdens_vo(Xv)
        n1v = fc.n1
    with lib.Operator(o - 2, v, sys_.naux, sys_.eps[keep], sys_.B[:, keep][:, :, keep]) as red:
        er, d2r, rhor = red.mp2()
    assert abs(e - er) < 1e-12 * abs(er) and abs(d2 - d2r) < 1e-12 * d2r
    emb = np.zeros_like(rho)
    emb[np.ix_(keep, keep)] = rhor
    emb[0, 0] = emb[2, 2] = 2.0
    ref_rho = mp2.density_frozen(sys_, [0, 2])
    assert np.abs(rho - emb).max() < 1e-12 and np.abs(rho - ref_rho).max() < 1e-12
    rows = np.array([a * o + i for a in range(v) for i in range(o) if i + 1 not in frozen])
    assert rows.size == n1v
    for s in range(2):
        x = np.zeros(o * v)
        x[rows] = Xv[:n1v, s]
        ref = nto.trdens_vo(sys_, x, rho=ref_rho)
        assert np.abs(T[s] - ref).max() < 1e-12 * np.abs(ref).max(), s
