"""Parity of the HIP sigma-vector, diagonal and ph-ph block against the oracle.

The oracle's literal restatement of the reference's element functions
(oracle/elements.py) is the checker at small sizes; at the benzene-sized
workload (o=16, v=200, naux=864, ndim 5.1e6) the density-fitted oracle
restatement (pinned against the literal one in test_oracle_pins) and
size-independent identities (symmetry <y,Hx> = <Hy,x>, linearity) are used.
Both coupling forms are checked: contractions with the density-fitting
factors ("df") and with the integral blocks assembled from them ("int").
FP64 tolerance: 1e-11 relative to max|sigma|; FP32 coupling 2e-6; BF16 3e-2.
"""
import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import dfsigma, elements

pytestmark = pytest.mark.gpu

# odd nocc and v a multiple of 8 (16-byte vector loads of bf16 rows) next to
# sizes that take the scalar GEMM loads
SMALL = [(2, 3, 2, 0), (3, 5, 7, 1), (4, 7, 11, 2), (5, 9, 13, 3), (5, 36, 164, 0), (6, 16, 40, 4), (3, 8, 10, 5),
         (7, 24, 30, 6)]
FORMS = ("df", "int")


@pytest.fixture(scope="module", params=[(c, f) for c in SMALL for f in FORMS],
                ids=lambda p: "o%d_v%d_naux%d-%s" % (p[0][:3] + (p[1],)))
def pair(request):
    (o, v, naux, seed), form = request.param
    print(f"[case] sigma o={o} v={v} naux={naux} seed={seed} coupling={form}", flush=True)
    sys = make_system(o, v, naux, seed=seed)
    blk = elements.build_blocks(sys, apply_minc=False)
    op = lib.Operator.from_system(sys, coupling=form)
    assert op.coupling == form
    yield sys, blk, op
    op.close()


def test_dims(pair):
    sys, blk, op = pair
    assert op.ndim == blk.ndim and op.n1 == blk.n1


def test_phph_block(pair):
    sys, blk, op = pair
    H11 = op.phph()
    scale = np.abs(blk.H11).max()
    assert np.abs(H11 - blk.H11).max() < 1e-12 * scale


def test_diagonal(pair):
    sys, blk, op = pair
    d = op.diagonal()
    assert np.abs(d - blk.diag()).max() < 1e-12 * np.abs(blk.diag()).max()


@pytest.mark.parametrize("nvec", [1, 3, 70])
def test_sigma_fp64(pair, nvec):
    sys, blk, op = pair
    rng = np.random.default_rng(nvec)
    X = rng.standard_normal((op.ndim, nvec))
    W = op.sigma(X)
    ref = blk.sigma(X)
    assert np.abs(W - ref).max() < 1e-11 * np.abs(ref).max()


@pytest.mark.parametrize("prec,tol", [("fp32", 2e-6), ("bf16", 3e-2)])
def test_sigma_lowprec(pair, prec, tol):
    sys, blk, op = pair
    rng = np.random.default_rng(5)
    X = rng.standard_normal((op.ndim, 4))
    W = op.sigma(X, precision=prec)
    ref = blk.sigma(X)
    assert np.abs(W - ref).max() < tol * np.abs(ref).max()


def test_sigma_unit_vectors_are_columns(pair):
    """sigma of unit vectors reproduces columns of the dense matrix (covers
    every configuration type, including the t4/t5 twins)."""
    sys, blk, op = pair
    H = blk.dense()
    n = op.ndim
    cols = np.unique(np.r_[0, blk.n1 - 1, blk.n1, blk.n1 + np.array(blk.type_offsets[1:-1]),
                           n - 1, np.linspace(0, n - 1, 11).astype(int)])
    cols = cols[cols < n]
    X = np.zeros((n, cols.size))
    X[cols, np.arange(cols.size)] = 1.0
    W = op.sigma(X)
    assert np.abs(W - H[:, cols]).max() < 1e-11 * np.abs(H).max()


@pytest.fixture(scope="module")
def benzene_sys():
    return make_system(16, 200, 864, seed=0)


@pytest.fixture(scope="module", params=FORMS)
def benzene(request, benzene_sys):
    print(f"[case] sigma benzene-size o=16 v=200 naux=864 coupling={request.param}", flush=True)
    op = lib.Operator.from_system(benzene_sys, coupling=request.param)
    yield benzene_sys, op
    op.close()


def test_benzene_auto_form_is_int(benzene_sys):
    """At o=16, v=200, naux=864 the integral form needs (v+o)/naux = 1/4 of
    the flops and fits easily, so the automatic choice takes it."""
    with lib.Operator.from_system(benzene_sys, second_order=False) as op:
        assert op.coupling == "int"


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16", 3e-2)])
def test_sigma_benzene_forms_agree(benzene_sys, prec, tol):
    """The two coupling forms evaluate the same matrix: FP64 to rounding,
    FP32/BF16 each within its tolerance of the other's FP64 result."""
    ops = {f: lib.Operator.from_system(benzene_sys, second_order=False, coupling=f) for f in FORMS}
    try:
        X = np.random.default_rng(21).standard_normal((ops["df"].ndim, 20))
        out = {f: (op.sigma(X), op.sigma(X, precision=prec)) for f, op in ops.items()}
    finally:
        for op in ops.values():
            op.close()
    ref = out["df"][0]
    scale = np.abs(ref).max()
    assert np.abs(out["int"][0] - ref).max() < 1e-12 * scale
    for form in FORMS:
        assert np.abs(out[form][1] - ref).max() < tol * scale


def test_sigma_benzene_symmetry_linearity(benzene):
    sys, op = benzene
    rng = np.random.default_rng(11)
    X = rng.standard_normal((op.ndim, 2))
    W = op.sigma(X)
    a = X[:, 1] @ W[:, 0]
    b = X[:, 0] @ W[:, 1]
    assert abs(a - b) < 1e-11 * np.abs(W).max() * np.sqrt(op.ndim) * 10
    W2 = op.sigma(2.0 * X[:, 0] - 3.0 * X[:, 1])
    assert np.abs(W2 - (2.0 * W[:, 0] - 3.0 * W[:, 1])).max() < 1e-11 * np.abs(W).max() * 5


def test_sigma_benzene_vs_df_oracle(benzene):
    sys, op = benzene
    df = dfsigma.DF(sys)
    H11 = op.phph()
    assert np.abs(H11 - df.H11).max() < 1e-11 * np.abs(df.H11).max()
    rng = np.random.default_rng(12)
    x = rng.standard_normal(op.ndim)
    w = op.sigma(x)
    ref = df.sigma(x)
    assert np.abs(w - ref).max() < 1e-11 * np.abs(ref).max()
    assert np.abs(op.diagonal() - df.diag()).max() < 1e-11 * np.abs(df.diag()).max()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("bf16", 3e-2)])
def test_sigma_many_occupied(prec, tol, form):
    """nocc > 64 takes the element-wise CSF<->pair kernels instead of the
    LDS-tiled ones; checked against the density-fitted oracle."""
    print(f"[case] sigma o=66 v=4 naux=30 {prec} coupling={form}", flush=True)
    sys = make_system(66, 4, 30, seed=6)
    df = dfsigma.DF(sys)
    with lib.Operator.from_system(sys, coupling=form) as op:
        rng = np.random.default_rng(13)
        X = rng.standard_normal((op.ndim, 3))
        W = op.sigma(X, precision=prec)
    ref = df.sigma(X)
    assert np.abs(W - ref).max() < tol * np.abs(ref).max()


@pytest.mark.parametrize("form", FORMS)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-11), ("fp32", 2e-6), ("bf16", 3e-2)])
def test_sigma_virtual_200(prec, tol, form):
    """v = 200 takes the 208-column GEMM tile of the INT down contraction
    (uneven split of the 13 column blocks over the waves); nocc = 2 keeps the
    density-fitted oracle fast.  9 vectors: a partial last vector group."""
    print(f"[case] sigma o=2 v=200 naux=20 {prec} coupling={form}", flush=True)
    sys = make_system(2, 200, 20, seed=8)
    df = dfsigma.DF(sys)
    with lib.Operator.from_system(sys, coupling=form) as op:
        X = np.random.default_rng(14).standard_normal((op.ndim, 9))
        W = op.sigma(X, precision=prec)
    ref = df.sigma(X)
    assert np.abs(W - ref).max() < tol * np.abs(ref).max()
