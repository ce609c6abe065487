"""Train the learned Davidson preconditioner on the GPU.

    python -m adc_ottawa_amd.train_precon [--base coupled] [--systems 200] [--epochs 30] [--out weights.npy]

The network multiplies a base correction step element by element by
(1 + m(features)): --base dpr is the diagonal (DPR) step of the reference,
--base coupled the coupled (Schur-complement) step.  Each base has its own
weights file (precon_weights.npy, precon_weights_coupled.npy).

1. Harvest: for each seeded synthetic system (o, v drawn from the given
   ranges, naux = 4 (o + v), seeds 1000, 1001, ...) run the FP64 Davidson with
   the base step and deflation twice -- once to convergence for the
   eigenvectors x*, once replaying the same trajectory -- and keep, for every
   unconverged root (matched to the x* it overlaps most) and iteration, 16
   sampled elements (half 1h1p, half 2h2p) of (features, base step, exact
   correction x - x*/(x.x*)), divided by the RMS element of the correction;
   at most --per-system records per system.  Records stay in HBM on one data
   handle.
2. Train: Adam on mini-batches with the library's own forward/backward
   kernels (adc2g_precon_train).
3. Save the fp32 master weights (4865 values) as .npy.

    python -m adc_ottawa_amd.train_precon --base coupled --evaluate 5000,5001

compares the saved network with its base step (and DPR) on unseen
benzene-sized systems.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from adc_ottawa_amd import lib  # noqa: E402
from adc_ottawa_amd.synth import make_system  # noqa: E402

WEIGHTS = {"dpr": os.path.join(HERE, "precon_weights.npy"),
           "coupled": os.path.join(HERE, "precon_weights_coupled.npy")}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def harvest(nsys, orange, vrange, nroots, seed0=1000, max_per_system=60000, base="dpr"):
    rng = np.random.default_rng(12345)
    data = None
    total = 0
    for i in range(nsys):
        o = int(rng.integers(orange[0], orange[1] + 1))
        v = int(rng.integers(vrange[0], vrange[1] + 1))
        seed = seed0 + i
        s = make_system(o, v, 4 * (o + v), seed=seed)
        op = lib.Operator.from_system(s)
        t0 = time.perf_counter()
        try:
            added = op.precon_harvest(nroots, max_records=max_per_system, base=base, blocksize=nroots,
                                      maxvec=10 * nroots, tol=1e-7, maxiter=600)
        except lib.Adc2gError as e:
            log(f"system {i} (o={o} v={v} seed={seed}) skipped: {e}")
            op.close()
            continue
        total += added
        log(f"system {i}: o={o} v={v} ndim={op.ndim} seed={seed}: {added} records "
            f"({time.perf_counter() - t0:.1f}s), total {total}")
        if data is None:
            data = op
        else:
            # move this system's records onto the data handle: train from each
            # handle in turn is equivalent, but one pool keeps batches mixed
            data.absorb(op)
            op.close()
    return data, total


def evaluate(weights, seeds, nroots=20, base="dpr"):
    """Time to nroots roots (1e-7, block nroots, subspace 10*nroots, deflation)
    with the base step alone and with the network on it, FP64 and with the
    low-precision first phase, on the benzene-sized system (o=16, v=200,
    naux=864) of each seed; DPR as the reference's step."""
    import json
    learned = "learned" if base == "dpr" else "learned_coupled"
    runs = [("fp64_dpr", dict(precon="dpr"))]
    if base == "coupled":
        runs += [("fp64_coupled", dict(precon="coupled")),
                 ("fp32_coupled_refined", dict(precon="coupled", precision="fp32", lowprec_tol=1e-4))]
    runs += [("fp64_" + learned, dict(precon=learned)),
             ("fp32_" + learned + "_refined", dict(precon=learned, precision="fp32", lowprec_tol=1e-4))]
    for seed in seeds:
        s = make_system(16, 200, 864, seed=seed)
        with lib.Operator.from_system(s) as op:
            op.set_precon_params(weights)
            ref = None
            for name, kw in runs:
                t0 = time.perf_counter()
                ev, _, st = op.davidson(nroots, blocksize=nroots, maxvec=10 * nroots, tol=1e-7, maxiter=1500,
                                        deflate=True, vectors=False, **kw)
                if ref is None:
                    ref = ev
                print(json.dumps(dict(seed=seed, run=name, seconds=round(time.perf_counter() - t0, 3),
                                      iterations=st["niter"], sigma_vectors=st["nsigma"],
                                      fallbacks=st["nfallback"], max_err_vs_fp64_dpr=float(abs(ev - ref).max()))),
                      flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--base", choices=("dpr", "coupled"), default="coupled")
    ap.add_argument("--systems", type=int, default=200)
    ap.add_argument("--omin", type=int, default=8)
    ap.add_argument("--omax", type=int, default=16)
    ap.add_argument("--vmin", type=int, default=80)
    ap.add_argument("--vmax", type=int, default=200)
    ap.add_argument("--roots", type=int, default=20)
    ap.add_argument("--per-system", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default="", help="weights file (default: the base's file in the package)")
    ap.add_argument("--evaluate", default="",
                    help="only evaluate the weights in --out on these comma-separated seeds (o=16, v=200)")
    a = ap.parse_args()
    out = a.out or WEIGHTS[a.base]
    if a.evaluate:
        evaluate(np.load(out), [int(x) for x in a.evaluate.split(",")], base=a.base)
        return
    t0 = time.perf_counter()
    data, total = harvest(a.systems, (a.omin, a.omax), (a.vmin, a.vmax), a.roots, max_per_system=a.per_system,
                          base=a.base)
    log(f"harvested {total} records in {time.perf_counter() - t0:.1f}s")
    t1 = time.perf_counter()
    loss = data.precon_train(epochs=a.epochs, batch=a.batch, lr=a.lr, seed=7)
    log(f"trained {a.epochs} epochs in {time.perf_counter() - t1:.1f}s; loss per epoch: "
        + " ".join(f"{x:.4f}" for x in loss))
    w = data.precon_params()
    np.save(out, w)
    log(f"saved {w.size} parameters to {out}")
    data.close()


if __name__ == "__main__":
    main()
