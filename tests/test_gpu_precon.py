"""Learned preconditioner: GPU forward/backward vs the CPU oracle, training,
and Davidson with the network.

Forward and gradients: the GPU evaluates in fp32 with the bf16-rounded
weights; the oracle (oracle/mlp.py, pinned by finite differences) in float64
with the same weights.  Bar: outputs within 1e-4 absolute, every gradient
within 2e-4 of the largest gradient component.
"""
import os

import numpy as np
import pytest

from adc_ottawa_amd import lib
from adc_ottawa_amd.synth import make_system
from oracle import dfsigma, mlp

pytestmark = pytest.mark.gpu

ORDER = ["W1", "E", "b1", "W2", "b2", "w3", "b3"]


def flat(p):
    return np.concatenate([p[k].reshape(-1) for k in ORDER]).astype(np.float32)


def unflat(v):
    shapes = {"W1": (64, 3), "E": (6, 64), "b1": (64,), "W2": (64, 64), "b2": (64,), "w3": (64,), "b3": (1,)}
    out, o = {}, 0
    for k in ORDER:
        n = int(np.prod(shapes[k]))
        out[k] = v[o:o + n].reshape(shapes[k]).astype(np.float64)
        o += n
    return out


@pytest.fixture(scope="module")
def op():
    print("[case] precon small o=3 v=8 naux=10", flush=True)
    sys = make_system(3, 8, 10, seed=4)
    o = lib.Operator.from_system(sys)
    yield sys, o
    o.close()


def test_forward_backward_vs_oracle(op):
    sys, o = op
    rng = np.random.default_rng(7)
    p = mlp.init_params(3)
    p["w3"] = mlp.bf16_round(rng.normal(0, 0.2, 64)).astype(np.float64)
    p["b3"] = np.array([0.125])
    o.set_precon_params(flat(p))
    assert o.precon_params().size == 4865
    n = 300
    x = mlp.transform(rng.normal(0, 3, n), rng.normal(1, 0.3, n), rng.normal(0, 1, n))
    cls = rng.integers(0, 6, n)
    dn = rng.normal(0, 1, n)
    en = rng.normal(0, 1, n)
    m_gpu, g_gpu, loss_gpu = o.precon_eval(x.astype(np.float32), cls, dn, en)
    xf = x.astype(np.float32).astype(np.float64)
    dnf = dn.astype(np.float32).astype(np.float64)
    enf = en.astype(np.float32).astype(np.float64)
    m_ref, _ = mlp.forward(p, xf, cls)
    assert np.abs(m_gpu - m_ref).max() < 1e-4
    assert abs(loss_gpu - mlp.loss(p, xf, cls, dnf, enf)) < 1e-4 * max(1.0, loss_gpu)
    g_ref = flat({k: v for k, v in mlp.backward(p, xf, cls, dnf, enf).items()})
    scale = np.abs(g_ref).max()
    assert np.abs(g_gpu - g_ref).max() < 2e-4 * scale


def test_harvest_train_and_solve(op):
    sys, o = op
    added = o.precon_harvest(4, tol=1e-8, maxiter=300)
    assert added > 0
    loss = o.precon_train(epochs=6, batch=512, lr=2e-3, seed=1)
    print("training loss per epoch:", loss)
    assert np.all(np.isfinite(loss))
    assert loss[-1] < loss[0]
    df = dfsigma.DF(sys)
    w = np.linalg.eigvalsh(df.sigma(np.eye(o.ndim)))[:4]
    ev, X, st = o.davidson(4, tol=1e-8, precon="learned", maxiter=400)
    assert np.abs(ev - w).max() < 1e-10
    r = np.linalg.norm(df.sigma(X) - X * ev, axis=0)
    assert r.max() < 1e-7


def test_learned_without_network_is_an_error():
    sys = make_system(3, 5, 6, seed=2)
    with lib.Operator.from_system(sys) as o:
        with pytest.raises(lib.Adc2gError) as e:
            o.davidson(2, precon="learned")
        assert e.value.code == 15


WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "adc_ottawa_amd",
                       "precon_weights.npy")


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained weights committed")
def test_trained_network_vs_dpr():
    """What the committed network gains over the DPR step on an unseen system.

    Both runs must converge to the same roots; the iteration ratio is printed.
    On these synthetic systems the trained rescaling of the DPR step does not
    pay off (the measured ratio is below 1); the sticky DPR fallback on
    stalled roots bounds the loss, which is what is asserted."""
    w = np.load(WEIGHTS)
    print("[case] learned vs dpr o=8 v=60 naux=272 seed=9001", flush=True)
    sys = make_system(8, 60, 272, seed=9001)
    with lib.Operator.from_system(sys) as o:
        o.set_precon_params(w)
        kw = dict(blocksize=10, maxvec=100, tol=1e-7, maxiter=1000, deflate=True, vectors=False)
        ev_d, _, st_d = o.davidson(10, precon="dpr", **kw)
        ev_l, _, st_l = o.davidson(10, precon="learned", **kw)
    gain = st_d["niter"] / st_l["niter"]
    print(f"DPR {st_d['niter']} iterations ({st_d['time_total']:.2f}s), learned {st_l['niter']} iterations "
          f"({st_l['time_total']:.2f}s, {st_l['nfallback']} DPR fallbacks): iteration gain {gain:.2f}")
    assert np.abs(ev_d - ev_l).max() < 1e-8
    assert gain >= 0.75


WEIGHTS_COUPLED = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "adc_ottawa_amd",
                               "precon_weights_coupled.npy")


@pytest.mark.skipif(not os.path.exists(WEIGHTS_COUPLED), reason="no trained weights committed")
def test_trained_coupled_network_benzene():
    """The committed network on the coupled step, on an unseen o=16, v=200
    system (seed 5002, not among the training seeds 1000-1199) with 20 roots:
    the same eigenvalues as the plain FP64 DPR Davidson in at most a tenth of
    its iterations with the FP32 first phase refined in FP64, and at least
    0.8 times the iteration efficiency of its base step (the coupled step
    without the network; the sticky fallback bounds the loss).  The measured
    ratios, wall times included, are printed."""
    w = np.load(WEIGHTS_COUPLED)
    print("[case] learned on coupled vs DPR o=16 v=200 naux=864 seed=5002, 20 roots", flush=True)
    sys_ = make_system(16, 200, 864, seed=5002)
    kw = dict(blocksize=20, maxvec=200, tol=1e-7, maxiter=1000, deflate=True, vectors=False)
    with lib.Operator.from_system(sys_) as o:
        o.set_precon_params(w)
        ev_c, _, st_c = o.davidson(20, precon="coupled", **kw)
        ev_n, _, st_n = o.davidson(20, precon="learned_coupled", **kw)
        ev_l, _, st_l = o.davidson(20, precon="learned_coupled", precision="fp32", lowprec_tol=1e-4, **kw)
        ev_d, _, st_d = o.davidson(20, precon="dpr", **kw)
    speedup = st_d["time_total"] / st_l["time_total"]
    gain = st_c["niter"] / st_n["niter"]
    print(f"DPR {st_d['niter']} it {st_d['time_total']:.2f}s; coupled {st_c['niter']} it {st_c['time_total']:.2f}s; "
          f"learned on coupled {st_n['niter']} it {st_n['time_total']:.2f}s ({st_n['nfallback']} base-step "
          f"fallbacks); learned on coupled + FP32 first phase {st_l['niter']} it {st_l['time_total']:.2f}s: "
          f"{speedup:.2f}x DPR, iteration gain of the network over its base {gain:.2f}")
    for ev in (ev_c, ev_n, ev_l):
        assert np.abs(ev - ev_d).max() < 1e-8
    assert st_l["niter"] * 10 <= st_d["niter"]
    assert gain >= 0.8
