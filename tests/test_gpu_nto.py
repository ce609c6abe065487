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


def test_trdens_vo_rejects_views_and_bad_sizes():
    print("[case] trdens_vo on a frozen-core view", flush=True)
    sys_ = make_system(3, 5, 8, seed=3)
    with lib.Operator.from_system(sys_) as op:
        with op.frozen_core([1]) as fc:
            with pytest.raises(lib.Adc2gError):
                fc.trdens_vo(np.ones(fc.ndim))
        with pytest.raises(ValueError):
            op.trdens_vo(np.ones(op.ndim - 1))


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
