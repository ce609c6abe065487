"""Block Lanczos pseudospectrum (adc2g_block_lanczos) against the oracle.

Small systems, both coupling forms: eigenvalues of T, |T| elementwise (the
GPU factor has a positive diagonal where the reference's Householder QR
takes signs, so T differs by a diagonal +-1 similarity), |transition
moments| and the Ritz vectors up to sign against oracle/lanczos.py (pinned
in test_oracle_pins); FP64 tolerances 1e-10 relative for T and its
eigenvalues, 1e-8 for moments and vectors; FP32 sigma-vectors 1e-5 for the
eigenvalues.  At the benzene size (ndim 5.1e6, 20 x 5 Lanczos vectors):
orthonormality of the Lanczos vectors, the band structure of T and exact
block Gauss quadrature sum_n (f.r_n)^2 theta_n^k = f^T H^k f (k <= 3) with
f^T H^k f from repeated FP64 sigma-vectors.
"""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import elements, lanczos

pytestmark = pytest.mark.gpu

FORMS = ("df", "int")


def _start_block(n, n1, b, seed):
    """Unit vectors on the b 1h1p configurations with the largest |mu|
    (init_vec_unit with stvc_lbl chosen by the transition dipoles) and a
    synthetic dipole vector mu on the 1h1p block."""
    mu = np.zeros(n)
    mu[:n1] = np.random.default_rng(seed).normal(0.0, 0.1, n1)
    idx = np.argsort(-np.abs(mu[:n1]), kind="stable")[:b]
    q0 = np.zeros((n, b))
    q0[idx, np.arange(b)] = 1.0
    return q0, mu


def _match_up_to_sign(a, b):
    d = np.minimum(np.abs(a - b).max(axis=0), np.abs(a + b).max(axis=0))
    return d.max()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("o,v,naux,seed,b,m", [(3, 8, 10, 5, 3, 6), (4, 12, 20, 6, 4, 6)])
def test_lanczos_small_vs_oracle(o, v, naux, seed, b, m, form):
    print(f"[case] lanczos o={o} v={v} naux={naux} b={b} m={m} coupling={form}", flush=True)
    sys = make_system(o, v, naux, seed=seed)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    n = H.shape[0]
    n1 = o * v
    q0, mu = _start_block(n, n1, b, seed)
    ref = lanczos.block_lanczos(lambda X: H @ X, q0, m)
    tm_ref = lanczos.tranmom(ref["ritz"], mu)
    with lib.Operator.from_system(sys, coupling=form) as op:
        ev, ex, st = op.block_lanczos(q0, m, travec=mu, tmat=True, ritz=True, check_orth=True)
        ev32, _, _ = op.block_lanczos(q0, m, precision="fp32")
    scale = np.abs(ref["eigval"]).max()
    assert st["nvec"] == b * m and st["nsigma"] == b * m
    assert st["orth_loss"] < 1e-10
    assert np.abs(ev - ref["eigval"]).max() < 1e-10 * scale
    assert np.abs(np.abs(ex["tmat"]) - np.abs(ref["T"])).max() < 1e-10 * scale
    assert np.abs(np.abs(ex["tranmom"]) - np.abs(tm_ref)).max() < 1e-8 * np.linalg.norm(mu)
    assert _match_up_to_sign(ex["ritz"], ref["ritz"]) < 1e-8
    assert np.abs(ev32 - ref["eigval"]).max() < 1e-5 * scale


def test_lanczos_bad_arguments():
    sys = make_system(3, 8, 10, seed=5)
    with lib.Operator.from_system(sys) as op:
        n = op.ndim
        q0, _ = _start_block(n, op.n1, 2, 0)
        with pytest.raises(lib.Adc2gError):
            op.block_lanczos(2.0 * q0, 3)          # not orthonormal
        with pytest.raises(lib.Adc2gError):
            op.block_lanczos(q0, 0)                # no cycles
        with pytest.raises(lib.Adc2gError):
            op.block_lanczos(q0, n // 2 + 1)       # more vectors than ndim
        # the handle is still usable after the failures
        ev, _, _ = op.block_lanczos(q0, 3)
        assert np.all(np.isfinite(ev))


def test_lanczos_benzene_quadrature():
    print("[case] lanczos benzene-size o=16 v=200 naux=864 b=20 m=5", flush=True)
    sys = make_system(16, 200, 864, seed=0)
    b, m = 20, 5
    with lib.Operator.from_system(sys) as op:
        n = op.ndim
        q0, _ = _start_block(n, op.n1, b, 7)
        c = np.random.default_rng(8).normal(size=b)
        f = q0 @ c
        ev, ex, st = op.block_lanczos(q0, m, travec=f, tmat=True, check_orth=True)
        hk = f.copy()
        moments = []
        for _ in range(4):
            moments.append(f @ hk)
            hk = op.sigma(hk)
    print(f"[lanczos benzene] orth_loss={st['orth_loss']:.2e} sigma {st['time_sigma']:.3f}s "
          f"T-eig {st['time_tdiag']:.3f}s total {st['time_total']:.3f}s", flush=True)
    assert st["orth_loss"] < 1e-9
    T = ex["tmat"]
    nv = b * m
    i, j = np.indices((nv, nv))
    assert np.abs(T[np.abs(i - j) > b]).max() == 0.0
    assert np.abs(T - T.T).max() == 0.0
    tm = ex["tranmom"]
    for k, ref in enumerate(moments):
        got = (tm ** 2 * ev ** k).sum()
        assert abs(got - ref) < 1e-9 * (tm ** 2 * np.abs(ev) ** k).sum(), k


@pytest.mark.parametrize("form", FORMS)
def test_lanczos_partial_reorthogonalisation_vs_oracle(form):
    """orthotype 1 on a run that loses orthogonality without it: the same
    number of re-orthogonalisations as the oracle, orthogonal stored vectors
    and every eigenvalue of T to 1e-10 relative (the extreme ones equal to
    the exact spectrum, the top one once); orthotype 0 on the same run loses
    orthogonality (max |L^T L - I| > 1e-2)."""
    print(f"[case] lanczos pro o=3 v=8 naux=10 b=3 m=40 coupling={form}", flush=True)
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    n = H.shape[0]
    q0 = np.zeros((n, 3))
    q0[[0, 5, 11], [0, 1, 2]] = 1.0
    ref = lanczos.block_lanczos(lambda X: H @ X, q0, 40, orthotype=1)
    lam = np.linalg.eigvalsh(H)
    with lib.Operator.from_system(sys, coupling=form) as op:
        ev, _, st = op.block_lanczos(q0, 40, orthotype=1, check_orth=True)
        ev0, _, st0 = op.block_lanczos(q0, 40, orthotype=0, check_orth=True)
    scale = np.abs(lam).max()
    assert st["npro"] == ref["npro"] > 0
    assert st["orth_loss"] < 1e-10 and st0["orth_loss"] > 1e-2
    assert np.abs(ev - ref["eigval"]).max() < 1e-10 * scale
    assert np.abs(ev[-3:] - lam[-3:]).max() < 1e-10 * scale
    # with re-orthogonalisation the top eigenvalue appears once (no ghost copy)
    assert (np.abs(ev - lam[-1]) < 1e-6).sum() == 1
    assert ev0.size == ev.size


@pytest.mark.parametrize("form", FORMS)
def test_lanczos_modified_partial_reorthogonalisation_vs_oracle(form):
    """orthotype 2 (mpro) on the run that loses orthogonality without
    re-orthogonalisation: the oracle's number of re-orthogonalisations,
    orthogonal stored vectors, every eigenvalue of T to 1e-10 relative and
    the top eigenvalue once."""
    print(f"[case] lanczos mpro o=3 v=8 naux=10 b=3 m=40 coupling={form}", flush=True)
    sys = make_system(3, 8, 10, seed=5)
    H = elements.build_blocks(sys, apply_minc=False).dense()
    n = H.shape[0]
    q0 = np.zeros((n, 3))
    q0[[0, 5, 11], [0, 1, 2]] = 1.0
    ref = lanczos.block_lanczos(lambda X: H @ X, q0, 40, orthotype=2)
    lam = np.linalg.eigvalsh(H)
    with lib.Operator.from_system(sys, coupling=form) as op:
        ev, _, st = op.block_lanczos(q0, 40, orthotype=2, check_orth=True)
    scale = np.abs(lam).max()
    assert st["npro"] == ref["npro"] > 0
    assert st["orth_loss"] < 1e-10
    assert np.abs(ev - ref["eigval"]).max() < 1e-10 * scale
    assert (np.abs(ev - lam[-1]) < 1e-6).sum() == 1


@pytest.mark.parametrize("kind,b", [(8, 2), (8, 4), (9, 3)])
def test_lanczos_guess_vs_oracle(kind, b):
    """init_vec_doublesubspace (8) / init_vec_singlesubspace (9) with a
    20-IS subspace: the oracle's columns up to sign, orthonormal."""
    print(f"[case] lanczos guess kind={kind} b={b} o=3 v=8 naux=10", flush=True)
    sys = make_system(3, 8, 10, seed=5)
    Hm = elements.build_blocks(sys, apply_minc=True).dense()
    ref = lanczos.subspace_guess(np.diag(Hm), lambda s: Hm[np.ix_(s, s)], b, kind=kind, subdim=20)
    with lib.Operator.from_system(sys) as op:
        q0 = op.lanczos_guess(kind, b, subdim=20)
        with pytest.raises(lib.Adc2gError):
            op.lanczos_guess(8, 3, subdim=20)         # odd block for the double subspace
    assert _match_up_to_sign(q0, ref) < 1e-10
    assert np.abs(q0.T @ q0 - np.eye(b)).max() < 1e-12


def test_spectral_bounds_lanczos_water_vs_oracle():
    """spectral_bounds 'lanczos' at the water size (ndim 16470, so the two
    800-IS subspaces are disjoint): the oracle's guess and 50-cycle block
    Lanczos (orthotype 0) give the same lowest and highest eigenvalues of T to
    1e-9 relative."""
    from oracle import dfsigma
    print("[case] spectral bounds lanczos water o=5 v=36 naux=164", flush=True)
    sys = make_system(5, 36, 164, seed=0)
    df = dfsigma.DF(sys)
    n = df.n1 + df.D2.size
    hii = df.diag()

    def hsub(sel):
        X = np.zeros((n, sel.size))
        X[sel, np.arange(sel.size)] = 1.0
        return df.sigma(X)[sel]

    q0 = lanczos.subspace_guess(hii, hsub, 2, kind=8, subdim=800)
    ref = lanczos.block_lanczos(df.sigma, q0, 50)
    with lib.Operator.from_system(sys) as op:
        lo, hi = op.spectral_bounds("lanczos")
    scale = np.abs(ref["eigval"]).max()
    print(f"[bounds lanczos] GPU {lo:.12f} {hi:.12f} oracle {ref['eigval'][0]:.12f} {ref['eigval'][-1]:.12f}",
          flush=True)
    assert abs(lo - ref["eigval"][0]) < 1e-9 * scale
    assert abs(hi - ref["eigval"][-1]) < 1e-9 * scale
