"""ADC(2) F-vector (adc2g_fvector) against the oracle.

Small systems, both coupling forms, symmetric and non-symmetric operator
matrices: the whole vector against the literal restatement of
get_modifiedtm_adc2 (oracle/fvec.py, pinned in test_oracle_pins) to 1e-12
relative (FP64 throughout).  At the water size (o=5, v=36, naux=164) against
the compact form of the oracle.  At the benzene size (o=16, v=200, naux=864;
ndim 5.1e6), where the oracle would need the nbas^4 integrals: the number
operator (dpl = identity) gives f = 0 to 1e-12 of a random operator's
vector and f is linear in dpl.  (The synthetic benzene-size system is
strongly correlated, E_MP2 ~ -2.7e4 and D2 ~ 20, so the second-order terms
outweigh the zeroth-order -sqrt(2) dpl(a,k) there.)
"""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import fvec

pytestmark = pytest.mark.gpu


def _dpl(nb, seed, symmetric=True):
    d = np.random.default_rng(seed).normal(0.0, 0.3, (nb, nb))
    return d + d.T if symmetric else d


@pytest.mark.parametrize("form", ("df", "int"))
@pytest.mark.parametrize("o,v,naux,seed", [(3, 8, 10, 5), (4, 12, 20, 6)])
def test_fvector_small_vs_oracle(o, v, naux, seed, form):
    print(f"[case] fvector o={o} v={v} naux={naux} coupling={form}", flush=True)
    sys = make_system(o, v, naux, seed=seed)
    nb = o + v
    with lib.Operator.from_system(sys, coupling=form) as op:
        for sym in (True, False):
            d = _dpl(nb, seed + 1, sym)
            got = op.fvector(d)
            ref = fvec.fvector(sys, d)
            assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max(), sym


def test_fvector_water_size_vs_compact_oracle():
    print("[case] fvector water size o=5 v=36 naux=164", flush=True)
    sys = make_system(5, 36, 164, seed=0)
    d = _dpl(41, 3)
    with lib.Operator.from_system(sys) as op:
        got = op.fvector(d)
    ref = fvec.fvector_compact(sys, d)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_fvector_benzene_number_operator_and_linearity():
    print("[case] fvector benzene size o=16 v=200 naux=864", flush=True)
    sys = make_system(16, 200, 864, seed=0)
    o, v = 16, 200
    nb = o + v
    d1, d2 = _dpl(nb, 11), _dpl(nb, 12)
    with lib.Operator.from_system(sys) as op:
        f1 = op.fvector(d1)
        f2 = op.fvector(d2)
        f12 = op.fvector(d1 + 2.0 * d2)
        f0 = op.fvector(np.eye(nb))
    scale = np.abs(f1).max()
    print(f"[fvector benzene] |f| {scale:.3e}, |f(I)| {np.abs(f0).max():.3e}", flush=True)
    assert np.all(np.isfinite(f1))
    assert np.abs(f0).max() < 1e-12 * scale
    assert np.abs(f12 - f1 - 2.0 * f2).max() < 1e-12 * np.abs(f12).max()


def test_fvector_bad_arguments():
    sys = make_system(3, 8, 10, seed=5)
    with lib.Operator.from_system(sys) as op:
        with pytest.raises(ValueError):
            op.fvector(np.eye(5))
        f = op.fvector(np.eye(11))
        assert np.abs(f).max() < 1e-14


@pytest.mark.parametrize("core", ([1], [1, 3]))
def test_fvector_cvs_view_vs_oracle(core):
    """On a CVS view the F-vector holds the CVS configurations' elements
    (get_modifiedtm_adc2 with lcvs, whose space has no I/II blocks): the
    oracle's full vector at oracle/csf.py's CVS indices, to 1e-12."""
    from oracle import csf
    print(f"[case] fvector cvs core={core} o=3 v=8 naux=10", flush=True)
    sys = make_system(3, 8, 10, seed=5)
    d = _dpl(11, 9)
    ref = fvec.fvector(sys, d)[csf.cvs_indices(3, 8, [c - 1 for c in core])]
    with lib.Operator.from_system(sys) as op:
        with op.cvs(core) as view:
            got = view.fvector(d)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()
