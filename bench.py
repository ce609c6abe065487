#!/usr/bin/env python3
"""Benchmark of the ADC(2)-s sigma-vector hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload benzene]

A step is one sigma-vector block product W = H X of the benzene-sized
synthetic system (o=16, v=200, naux=864, ndim 5,124,800) on a block of 20
resident FP64 vectors (the Davidson block size of that workload).  Inputs are
in HBM before the timed region; the JSON line reports the whole-job
sigma-vectors/s, the roofline of the dominant kernel (the virtual-index
coupling contractions, FP64 MFMA; the naux contractions with --coupling df)
measured with HIP events over the timed region, and a CPU baseline: the reference's stored-matrix sigma-vector
(COO + hxvec_incore) restated in C/OpenMP and timed on the host on a sample of
matrix rows, scaled to the full matrix.

After the timed steps it also runs the solver comparison the metric names:
time to converge 20 roots to a residual of 1e-7 with the plain FP64 Davidson
(diagonal preconditioner), the mixed-precision Davidson, and the learned
preconditioner when trained weights are present.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from adc_ottawa_amd import lib  # noqa: E402
from adc_ottawa_amd.synth import SIZES, make_system  # noqa: E402

METRIC = "sigma-vector GB/s & time-to-20-roots (≤1e-7 Eh): learned vs diagonal precond"
PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0}  # dense MFMA TFLOP/s, MI355X
HBM_PEAK = 8000.0  # GB/s
WEIGHTS = os.path.join(ROOT, "adc_ottawa_amd", "precon_weights.npy")
WEIGHTS_COUPLED = os.path.join(ROOT, "adc_ottawa_amd", "precon_weights_coupled.npy")
# HBM bytes per launch of the coupling contractions, measured with rocprofv3
# FETCH_SIZE / WRITE_SIZE counter runs of this benchmark (scaled by a copy of
# known size); written next to the rocprof summaries it comes from.
TRAFFIC = os.path.join(ROOT, "profiles", "coupling_gemm_traffic.json")


def measured_traffic(prec, workload, form):
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
        rec = t.get(workload, {}).get(form, {}).get(prec)
        return None if rec is None else float(rec["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime binding for device buffers (no torch needed)."""

    def __init__(self):
        self.h = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]

    def malloc(self, nbytes):
        p = C.c_void_p()
        rc = self.h.hipMalloc(C.byref(p), nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc({nbytes}) failed: {rc}")
        return p.value

    def h2d(self, dst, arr):
        rc = self.h.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy failed: {rc}")

    def sync(self):
        rc = self.h.hipDeviceSynchronize()
        if rc != 0:
            raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")

    def free(self, p):
        self.h.hipFree(C.c_void_p(p))

    def d2d(self, dst, src, nbytes):
        rc = self.h.hipMemcpy(C.c_void_p(dst), C.c_void_p(src), nbytes, 3)  # hipMemcpyDeviceToDevice
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2D failed: {rc}")


def copy_calibration(hip, nbytes=1 << 30, reps=3):
    """Plain device-to-device copies of a known size (runtime blit kernel).

    Under ``rocprofv3 --pmc FETCH_SIZE`` / ``WRITE_SIZE`` the counters read for
    these dispatches, against the nbytes read and written, give the factor that
    turns the counters of the sigma kernels into HBM bytes."""
    a, b = hip.malloc(nbytes), hip.malloc(nbytes)
    for _ in range(reps):
        hip.d2d(b, a, nbytes)
    hip.sync()
    hip.free(a)
    hip.free(b)
    log(f"copy calibration: {reps} device copies of {nbytes} bytes")


def timed_sigma(op, hip, dX, dW, nvec, prec, steps, warmup):
    for _ in range(warmup):
        op.sigma_device(nvec, dX, dW, prec)
    hip.sync()
    op.profile(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        op.sigma_device(nvec, dX, dW, prec)
    hip.sync()
    t1 = time.perf_counter()
    prof = op.profile(False)
    return (t1 - t0), prof


KERNEL_DESC = {
    "int": "gemm_kernel<%s,*,tag=1> (virtual-index coupling contractions with the (vv|ov) block: "
           "down + up, mean per launch)",
    "df": "gemm_kernel<%s,128,128,*,tag=1> (DF coupling contractions over naux: down + up, mean per launch)",
}


def roofline_of(prof, prec, workload, form):
    ms, launches, flops = prof["coupling_gemm"]
    achieved = flops / (ms * 1e-3) / 1e12
    return {"kernel": KERNEL_DESC[form] % prec,
            "bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK[prec], "unit": "TFLOP/s",
            "frac": round(achieved / PEAK[prec], 4), "traffic": measured_traffic(prec, workload, form),
            "launches": int(launches), "avg_launch_ms": round(ms / max(launches, 1), 4),
            "flops_per_launch": flops / max(launches, 1)}


def minor_stats(prof):
    ms, n, flops = prof["coupling_minor"]
    return {"ms": round(ms, 4), "launches": int(n), "TFLOPs": round(flops / max(ms, 1e-9) * 1e-9, 2)}


def stream_stats(prof):
    out = {}
    for k in ("csf_to_pair", "pair_to_csf_diag"):
        ms, n, b = prof[k]
        out[k] = {"GBps": round(b / (ms * 1e-3) / 1e9, 1), "frac_of_8TBps": round(b / (ms * 1e-3) / 1e9 / HBM_PEAK, 3),
                  "avg_launch_ms": round(ms / max(n, 1), 4), "bytes_per_launch": b / max(n, 1)}
    return out


def host_info():
    """The host the CPU timings ran on: logical CPUs and the model name."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"logical_cpus": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "model": model}


def cpu_time_to_roots(nroots=5):
    """Water-sized system (o=5, v=36, naux=164, the reference's CPU-runnable
    case): time to nroots roots (block nroots, subspace 4*block, residual
    1e-7, no deflation) of the reference's algorithm on the host -- the
    Davidson restatement (oracle/davidson.py: all currdim sigma-vectors every
    iteration, QR of the subspace) with DPR on the full stored matrix through
    the C/OpenMP hxvec_incore restatement -- beside the GPU Davidson on the
    same system with DPR, Olsen (the reference's default) and the coupled
    step."""
    from oracle import cpu_baseline as cb  # the checker / baseline, never the product
    from oracle import davidson as odav
    s0 = make_system(5, 36, 164, seed=0)
    out = {"system": "o=5 v=36 naux=164 seed 0", "roots": nroots, "host": host_info()}
    with lib.Operator.from_system(s0) as op0:
        h11, hii = op0.phph(), op0.diagonal()
        for name, pc in (("dpr", "dpr"), ("olsen", "olsen"), ("coupled", "coupled")):
            t0 = time.perf_counter()
            ev, _, st = op0.davidson(nroots, blocksize=nroots, maxvec=4 * nroots, tol=1e-7, precon=pc,
                                     deflate=False, maxiter=2000, vectors=False)
            out["gpu_" + name] = {"seconds": round(time.perf_counter() - t0, 4), "iterations": st["niter"],
                                  "sigma_vectors": st["nsigma"]}
            if name == "dpr":
                ev_gpu = ev
    coo = cb.COO(s0, h11=h11)
    sigma = lambda V: coo.hxvec(hii, V)[0]  # noqa: E731
    for name, pc in (("dpr", 1),):
        t0 = time.perf_counter()
        r = odav.davidson(sigma, hii, nroots, blocksize=nroots, maxvec=4 * nroots, tol=1e-7, precon=pc,
                          deflate=False, maxiter=2000)
        wall = time.perf_counter() - t0
        out["cpu_" + name] = {"seconds": round(wall, 3), "iterations": r["niter"], "sigma_vectors": r["nmult"],
                              "max_abs_err_vs_gpu_Eh": float(np.abs(r["eigval"] - ev_gpu).max())}
        log(f"cpu time-to-{nroots}-roots ({name}, water size): {wall:.2f}s, {r['niter']} iterations")
    return out


def cpu_baseline(sys_, op, nvec):
    from oracle import cpu_baseline as cb  # the checker / baseline, never the product
    cb.build()
    L = cb.load()
    o, v = sys_.nocc, sys_.nvirt
    per_row = L.coo_row_coupling_count(o, v)
    na = max(1, min(v, int(round(1.5e8 / (o * per_row)))))
    t0 = time.perf_counter()
    coo = cb.COO(sys_, h11=op.phph(), a0=0, a1=na)
    tb = time.perf_counter() - t0
    hii = op.diagonal()
    X = np.random.default_rng(1).standard_normal((coo.ndim, nvec))
    _, t_off, t_diag = coo.hxvec(hii, X)
    scale = coo.full_nnz / coo.nnz
    t_full = t_diag + t_off * scale
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    log(f"cpu baseline: sample {coo.rows} rows, nnz {coo.nnz} of {coo.full_nnz}, build {tb:.1f}s, "
        f"hxvec offdiag {t_off:.2f}s diag {t_diag:.2f}s -> full block {t_full:.1f}s")
    return {"value": round(nvec / t_full, 4), "unit": "sigma-vectors/s", "cores": min(cores, nvec),
            "kind": "port", "host": host_info(),
            "sample": (f"reference stored-matrix path restated in C/OpenMP (COO build as get_offdiag_adc2_save_omp, "
                       f"sigma as hxvec_incore, OpenMP over the {nvec} columns): COO of the {coo.rows} 1h1p rows "
                       f"with virtual index < {na} ({coo.nnz} of {coo.full_nnz} stored elements), hxvec_incore "
                       f"timed on {nvec} vectors ({t_off:.2f}s off-diagonal + {t_diag:.2f}s diagonal) and the "
                       f"off-diagonal time scaled by the stored-element ratio {scale:.1f}"),
            "seconds_per_block_scaled": round(t_full, 2), "stored_elements_full": int(coo.full_nnz)}


def davidson_compare(op, nroots, only=None, blocksize=20, maxvec=200):
    """Time to nroots roots (residual 1e-7, deflation) with the reference's
    steps (DPR; Olsen, the reference's default precon), the coupled
    (Schur-complement) step, the learned networks on DPR and on the coupled
    step, each in FP64 and with an FP32 first phase refined in FP64."""
    res = {}
    runs = [("fp64_dpr", dict(precon="dpr", precision="fp64")),
            ("fp64_olsen", dict(precon="olsen", precision="fp64")),
            ("fp32_dpr_refined", dict(precon="dpr", precision="fp32", lowprec_tol=1e-4)),
            ("fp64_coupled", dict(precon="coupled", precision="fp64")),
            ("fp32_coupled_refined", dict(precon="coupled", precision="fp32", lowprec_tol=1e-4))]
    have_net = os.path.exists(WEIGHTS)
    have_cnet = os.path.exists(WEIGHTS_COUPLED)
    if have_net:
        runs += [("fp64_learned", dict(precon="learned", precision="fp64", weights=WEIGHTS))]
    if have_cnet:
        runs += [("fp64_learned_coupled", dict(precon="learned_coupled", precision="fp64", weights=WEIGHTS_COUPLED)),
                 ("fp32_learned_coupled_refined", dict(precon="learned_coupled", precision="fp32",
                                                       lowprec_tol=1e-4, weights=WEIGHTS_COUPLED))]
    if only:
        runs = [r for r in runs if r[0] in only]
    ref = None
    for name, kw in runs:
        kw = dict(kw)
        w = kw.pop("weights", None)
        if w is not None:
            op.set_precon_params(np.load(w))
        t0 = time.perf_counter()
        try:
            ev, _, st = op.davidson(nroots, blocksize=blocksize, maxvec=maxvec, tol=1e-7, maxiter=1000,
                                    deflate=True, vectors=False, **kw)
        except lib.Adc2gError as e:
            res[name] = {"error": str(e)}
            log(f"davidson {name}: {e}")
            continue
        wall = time.perf_counter() - t0
        rec = {"seconds": round(wall, 3), "iterations": st["niter"], "sigma_vectors": st["nsigma"],
               "lowprec_iterations": st["niter_lowprec"], "fallbacks": st["nfallback"],
               "sigma_seconds": round(st["time_sigma"], 3), "max_resnorm": st["resnorm_max"]}
        if st["nprecon_builds"]:
            rec.update({"precon_seconds": round(st["time_precon"], 3), "schur_complements": st["nprecon_builds"],
                        "shift_Eh": round(st["precon_sigma"], 6)})
        if ref is None:
            ref = ev
            rec["eigenvalues_Eh"] = [float(x) for x in ev]
        else:
            err = np.abs(ev - ref)
            rec["max_abs_err_vs_fp64_Eh"] = float(err.max())
            rec["per_root_err_vs_fp64_Eh"] = [float(x) for x in err]
        res[name] = rec
        log(f"davidson {name}: {wall:.2f}s, {st['niter']} iterations, {st['nsigma']} sigma-vectors")
    base = res.get("fp64_dpr", {}).get("seconds")
    if base:
        for name in res:
            if "seconds" in res[name]:
                res[name]["speedup_vs_fp64_dpr"] = round(base / res[name]["seconds"], 3)
    res["learned_weights"] = {"dpr": have_net, "coupled": have_cnet}
    res["roots"], res["blocksize"], res["maxvec"] = nroots, blocksize, maxvec
    return res


def sil_compare(op, sz, tfinal, autotol, seed=0):
    """SIL propagation of the benzene-sized Hamiltonian: initial state the
    seeded synthetic dipole vector (1h1p part, N(0, 0.1^2)) normalised, Lanczos
    order 20, output interval 0.1 a.u., autocorrelation every interval; FP64
    and FP32-MFMA sigma-vectors (FP64 recursion and exponential in both)."""
    from adc_ottawa_amd.synth import unit_dipole
    f = np.zeros(op.ndim)
    f[: op.n1] = unit_dipole(sz["nocc"], sz["nvirt"], seed=seed)
    f /= np.linalg.norm(f)
    out, ref = {}, None
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        auto, st = op.sil_autocorr(f, tout=0.1, tfinal=tfinal, kdim=20, autotol=autotol, precision=prec,
                                   verbose=True)
        wall = time.perf_counter() - t0
        rec = {"seconds": round(wall, 3), "records": int(auto.size), "lanczos_steps": int(st["nsteps"]),
               "sigma_vectors": int(st["nsigma"]), "max_order": int(st["max_order"]),
               "ms_per_complex_sigma_vector": round(1e3 * wall / max(int(st["nsigma"]), 1), 4),
               "max_abs_autocorrelation_minus_1_at_t0": float(abs(auto[0] - 1.0))}
        if ref is None:
            ref = auto
        else:
            n = min(ref.size, auto.size)
            rec["max_abs_err_vs_fp64"] = float(np.abs(auto[:n] - ref[:n]).max())
        out[prec] = rec
        log(f"sil {prec}: {wall:.2f}s, {st['nsteps']} Lanczos steps, {st['nsigma']} sigma-vectors, "
            f"{auto.size} autocorrelation records")
    out["tfinal_au"] = tfinal
    out["autotol"] = autotol
    return out


def cheby_measure(op, sz, order, seed=0):
    """Chebyshev order-domain autocorrelation (adc2g_cheby_auto) of the
    synthetic dipole vector; FP64 and FP32 sigma-vectors.  Bounds: the
    library's spectral_bounds 'lanczos' (block Lanczos from the
    double-subspace guess, specbounds.f90:52-105) widened by 5 % of the range
    on each side (the reference's 0.9 / 1.1 widening assumes positive
    bounds); 'davlanc' is timed and reported beside it."""
    from adc_ottawa_amd.synth import unit_dipole
    t0 = time.perf_counter()
    lo, hi = op.spectral_bounds("lanczos")
    t_lz = time.perf_counter() - t0
    log(f"spectral bounds (lanczos): {lo:.6f} .. {hi:.3f} Eh in {t_lz:.2f}s")
    pad = 0.05 * (hi - lo)
    bounds = (lo - pad, hi + pad)
    t0 = time.perf_counter()
    try:
        dl = op.spectral_bounds("davlanc")
        log(f"spectral bounds (davlanc): {dl[0]:.6f} .. {dl[1]:.3f} Eh in {time.perf_counter() - t0:.2f}s")
    except lib.Adc2gError as e:   # reported; the widened Lanczos bounds are used either way
        dl = str(e)
        log(f"spectral bounds (davlanc): {e}")
    t_dl = time.perf_counter() - t0
    q0 = np.zeros(op.ndim)
    q0[: op.n1] = unit_dipole(sz["nocc"], sz["nvirt"], seed=seed)
    q0 /= np.linalg.norm(q0)
    out, ref = {"bounds_Eh": bounds, "lanczos_bounds_Eh": (lo, hi), "lanczos_bounds_seconds": round(t_lz, 3),
                "order": order, "davlanc_bounds_Eh": dl, "davlanc_seconds": round(t_dl, 3)}, None
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        auto = op.cheby_auto(q0, bounds, order, precision=prec)
        wall = time.perf_counter() - t0
        rec = {"seconds": round(wall, 3), "ms_per_order": round(wall / order * 1e3, 3)}
        if ref is None:
            ref = auto
        else:
            rec["max_abs_err_vs_fp64"] = float(np.abs(auto - ref).max())
        out[prec] = rec
        log(f"cheby {prec}: {wall:.2f}s for {order} orders ({2 * order + 1} autocorrelation values)")
    return out


def lanczos_measure(op, sz, spec, seed=0):
    """Block Lanczos pseudospectrum (adc2g_block_lanczos): blocksize x ncycles
    from unit vectors on the 1h1p configurations with the largest synthetic
    transition dipoles (init_vec_unit), transition moments with that dipole
    vector; FP64 and FP32 sigma-vectors."""
    from adc_ottawa_amd.synth import unit_dipole
    b, m = (int(x) for x in spec.split(","))
    mu = np.zeros(op.ndim)
    mu[: op.n1] = unit_dipole(sz["nocc"], sz["nvirt"], seed=seed)
    idx = np.argsort(-np.abs(mu[: op.n1]), kind="stable")[:b]
    q0 = np.zeros((op.ndim, b), order="F")
    q0[idx, np.arange(b)] = 1.0
    out, ref = {"blocksize": b, "ncycles": m, "nvec": b * m}, None
    for prec in ("fp64", "fp32"):
        t0 = time.perf_counter()
        ev, ex, st = op.block_lanczos(q0, m, precision=prec, travec=mu, check_orth=(prec == "fp64"))
        wall = time.perf_counter() - t0
        rec = {"seconds": round(wall, 3), "sigma_seconds": round(st["time_sigma"], 3),
               "tdiag_seconds": round(st["time_tdiag"], 3), "ritz_seconds": round(st["time_ritz"], 3),
               "lowest_ritz_Eh": [float(x) for x in ev[:3]]}
        if prec == "fp64":
            rec["orth_loss"] = float(st["orth_loss"])
            ref = ev
        else:
            rec["max_abs_ritz_diff_vs_fp64_lowest20"] = float(np.abs(ev[:20] - ref[:20]).max())
        out[prec] = rec
        log(f"lanczos {prec}: {wall:.2f}s for {b}x{m} (sigma {st['time_sigma']:.2f}s, "
            f"T eig {st['time_tdiag']:.2f}s, Ritz {st['time_ritz']:.2f}s)")
    return out


def relax_measure(op, tol):
    """Imaginary-time relaxation (adc2g_relax, XSIL, krydim 10, step 30 a.u.)
    of the lowest state from the unit vector on the smallest |H_jj|, FP64."""
    t0 = time.perf_counter()
    ev, _, res, nit, st = op.relax(1, step=30.0, maxiter=400, tol=tol, vectors=False)
    wall = time.perf_counter() - t0
    out = {"tol": tol, "step_au": 30.0, "krydim": 10, "seconds": round(wall, 3),
           "timesteps": int(nit[0]), "sigma_vectors": int(st["nsigma"]),
           "sigma_seconds": round(st["time_sigma"], 3),
           "ms_per_timestep": round(wall / max(1, int(nit[0])) * 1e3, 3),
           "energy_Eh": float(ev[0]), "residual": float(res[0])}
    log(f"relax fp64: {wall:.2f}s, {int(nit[0])} timesteps to residual {res[0]:.2e}, E = {ev[0]:.8f}")
    return out


def mp2_measure(op):
    """MP2 correlation energy, D2 diagnostic and ground-state density
    (adc2g_mp2) of the workload's system, FP64."""
    t0 = time.perf_counter()
    e, d2, _ = op.mp2(density=False)
    t1 = time.perf_counter()
    _, _, rho = op.mp2(density=True)
    t2 = time.perf_counter()
    log(f"mp2: energy + D2 {t1 - t0:.2f}s, with the density {t2 - t1:.2f}s")
    return {"energy_seconds": round(t1 - t0, 3), "with_density_seconds": round(t2 - t1, 3),
            "e_mp2_Eh": float(e), "d2": float(d2), "density_trace": float(np.trace(rho))}


def fvector_measure(op):
    """ADC(2) F-vector (adc2g_fvector) of a seeded symmetric one-electron
    operator matrix on the workload's system, FP64; the number operator
    (identity) must give f = 0."""
    nb = op.nocc + op.nvirt
    d = np.random.default_rng(7).normal(0.0, 0.3, (nb, nb))
    d = d + d.T
    op.fvector(np.eye(nb))  # first call: kernel loads
    t0 = time.perf_counter()
    f = op.fvector(d)
    t1 = time.perf_counter()
    f0 = op.fvector(np.eye(nb))
    log(f"fvector: {t1 - t0:.3f}s, |f| {np.linalg.norm(f):.4e}, |f(I)|max {np.abs(f0).max():.1e}")
    return {"seconds": round(t1 - t0, 4), "norm": float(np.linalg.norm(f)),
            "number_operator_max": float(np.abs(f0).max())}


def trdens_measure(op, nvec=20, seed=0):
    """Occupied-virtual transition density (adc2g_trdens_vo) of nvec seeded
    vectors: the first call builds M (n1 x n1, one DF GEMM) and the MP2
    density, later calls are two n1 x n1 x nvec GEMMs and one kernel."""
    X = np.asfortranarray(np.random.default_rng(seed).standard_normal((op.ndim, nvec)))
    t0 = time.perf_counter()
    op.trdens_vo(X[:, :1])
    t1 = time.perf_counter()
    T = op.trdens_vo(X)
    t2 = time.perf_counter()
    log(f"trdens: first call (M and the MP2 density) {t1 - t0:.2f}s, {nvec} vectors {t2 - t1:.3f}s")
    return {"first_call_seconds": round(t1 - t0, 3), "seconds": round(t2 - t1, 4), "nvec": nvec,
            "norm": float(np.linalg.norm(T))}


def fdstates_measure(op, sz, order, nsel=10, seed=0):
    """Chebyshev filter-diagonalisation states (adc2g_cheby_fdstates): order
    Chebyshev orders from the synthetic dipole vector, nsel states with
    seeded coefficients, FP64; bounds from the library's spectral_bounds
    'lanczos' widened by 5 % of the range."""
    from adc_ottawa_amd.synth import unit_dipole
    lo, hi = op.spectral_bounds("lanczos")
    pad = 0.05 * (hi - lo)
    q0 = np.zeros(op.ndim)
    q0[: op.n1] = unit_dipole(sz["nocc"], sz["nvirt"], seed=seed)
    coef = np.random.default_rng(seed).standard_normal((order + 1, nsel)) / (1.0 + np.arange(order + 1))[:, None]
    t0 = time.perf_counter()
    e, _ = op.cheby_fdstates(q0, (lo - pad, hi + pad), coef)
    wall = time.perf_counter() - t0
    log(f"cheby fdstates: {wall:.2f}s for {order} orders and {nsel} states")
    return {"seconds": round(wall, 3), "ms_per_order": round(wall / order * 1e3, 3), "order": order,
            "nsel": nsel, "energies_Eh": [float(x) for x in e]}


def laser_measure(op, tfinal, nnz=20_000_000):
    """Laser-driven propagation of the ground state (adc2g_sil_laser) under
    a cos^2 pulse along x, with a synthetic intermediate-state dipole of nnz
    off-diagonal elements (the reference computes it in its ISR code; its
    sparsity here is an assumption), then the same with a diagonal CAP
    (Lanczos-Arnoldi steps).  tout = 1 a.u."""
    rng = np.random.default_rng(5)
    nd = op.ndim
    row = rng.integers(0, nd, nnz)
    col = (row + 1 + rng.integers(0, nd - 1, nnz)) % nd
    op.set_dipole_is(0, rng.normal(0.0, 0.1, nd), row, col, rng.normal(0.0, 1e-3, nnz), 0.3,
                     rng.normal(0.0, 1e-3, nd))
    del row, col
    psi0 = np.zeros(nd + 1, dtype=complex)
    psi0[nd] = 1.0
    pulse = [{"envelope": "cos2", "envpar": (1.0, 1.0), "carrier": "cos", "freq": 0.5,
              "pulse_vec": (1.0, 0.0, 0.0), "strength": 0.05}]
    out = {"tfinal_au": tfinal, "dipole_offdiagonal_elements": nnz}
    for name, cap in (("sil", False), ("csil_cap", True)):
        if cap:
            op.set_cap_is(np.abs(rng.normal(0.0, 1e-2, nd)), [], [], [], 0.0, np.zeros(nd))
        t0 = time.perf_counter()
        psi, norms, st = op.sil_laser(psi0, pulse, tout=1.0, tfinal=tfinal, kdim=20, proptol=1e-8)
        dt = time.perf_counter() - t0
        out[name] = {"seconds": round(dt, 3), "steps": st["nsteps"], "operator_products": st["nsigma"],
                     "ms_per_product": round(1e3 * dt / max(st["nsigma"], 1), 3),
                     "final_norm": float(norms[-1]), "excited_norm": float(np.linalg.norm(psi[:nd]))}
        log(f"laser {name}: {tfinal} a.u. in {dt:.2f}s, {st['nsteps']} steps, {st['nsigma']} operator products, "
            f"norm {norms[-1]:.9f}")
    op.set_cap_is(None, [], [], [], 0.0, None)
    op.set_dipole_is(0, None, [], [], [], 0.0, None)
    return out


def adc2x_measure(op, nvec=20, roots=0):
    """ADC(2)-x on the workload's operator (adc2g_set_adc2x): the set-up (the
    (vv|vv), (oo|oo) and ring matrices and the diagonal), and FP64
    sigma-vectors of a block of nvec vectors with the first-order 2h2p-2h2p
    block; with roots > 0 also a Davidson (Olsen, block = roots, subspace 10
    blocks, 1e-7, deflation).  Switched off again afterwards."""
    rng = np.random.default_rng(3)
    X = np.asfortranarray(rng.standard_normal((op.ndim, nvec)))
    t0 = time.perf_counter()
    op.set_adc2x(True)
    t_setup = time.perf_counter() - t0
    try:
        op.sigma(X[:, :2])
        t0 = time.perf_counter()
        op.sigma(X)
        t_sig = time.perf_counter() - t0
        out = {"setup_seconds": round(t_setup, 3), "block": nvec, "ms_per_sigma_vector": round(1e3 * t_sig / nvec, 3)}
        log(f"adc2x: set-up {t_setup:.2f}s, {1e3 * t_sig / nvec:.2f} ms per sigma-vector")
        if roots > 0:
            t0 = time.perf_counter()
            ev, _, st = op.davidson(roots, blocksize=roots, maxvec=10 * roots, tol=1e-7, precon="olsen",
                                    deflate=True, maxiter=1000, vectors=False)
            out["davidson"] = {"roots": roots, "seconds": round(time.perf_counter() - t0, 2),
                               "iterations": st["niter"], "sigma_vectors": st["nsigma"],
                               "eigenvalues_Eh": [float(x) for x in ev]}
            log(f"adc2x davidson: {roots} roots in {out['davidson']['seconds']}s, {st['niter']} iterations")
    finally:
        op.set_adc2x(False)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="benzene", choices=sorted(SIZES))
    ap.add_argument("--nvec", type=int, default=20)
    ap.add_argument("--coupling", default="auto", choices=("auto", "df", "int"),
                    help="coupling form of the operator (auto: the library's choice)")
    ap.add_argument("--skip-davidson", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-lowprec", action="store_true")
    ap.add_argument("--sil", type=float, default=-1.0,
                    help="SIL propagation (Lanczos order 20, output every 0.1 a.u., FP64 and FP32 sigma) to this "
                         "final time in a.u.; default 10 (201 autocorrelation records, ~40 s) for the benzene "
                         "workload, 0 skips; the full run of 2001 records is --sil 100")
    ap.add_argument("--sil-tol", type=float, default=1e-10, help="SIL error tolerance (autotol)")
    ap.add_argument("--cheby", type=int, default=500,
                    help="Chebyshev order-domain autocorrelation to this order (0 skips)")
    ap.add_argument("--lanczos", default="50,20",
                    help="block Lanczos pseudospectrum, 'blocksize,ncycles' (empty skips)")
    ap.add_argument("--relax", type=float, default=1e-3,
                    help="relax the lowest state in imaginary time to this residual (0 skips)")
    ap.add_argument("--skip-mp2", action="store_true", help="skip the MP2 energy/density timing")
    ap.add_argument("--skip-fvector", action="store_true", help="skip the F-vector timing")
    ap.add_argument("--laser", type=float, default=2.0,
                    help="laser / CAP propagation leg over this many a.u. (0 skips)")
    ap.add_argument("--trdens", action="store_true", help="time the transition density of 20 vectors")
    ap.add_argument("--fdstates", type=int, default=0,
                    help="Chebyshev filter-diagonalisation states to this order (0 = skip)")
    ap.add_argument("--adc2x-roots", type=int, default=0,
                    help="ADC(2)-x leg: also a Davidson to this many roots (the sigma-vector timing runs by "
                         "default; -1 skips the leg)")
    ap.add_argument("--dav-roots", type=int, default=0,
                    help="Davidson comparison: roots (default 20; 40 for the large workload)")
    ap.add_argument("--dav-block", type=int, default=0,
                    help="Davidson comparison: block size, >= roots (default: the number of roots)")
    ap.add_argument("--dav-maxvec", type=int, default=0,
                    help="Davidson comparison: maximum subspace (default 10 blocks; 2 for the large workload)")
    ap.add_argument("--dav-runs", default="",
                    help="comma-separated subset of the Davidson comparison runs (default: all; for the large "
                         "workload the coupled step in FP64 and with the FP32 first phase)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run device copies of 1 GiB (scale factor for rocprofv3 --pmc byte counters)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the sigma-vector on a single MI355X "
              "(no collectives); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    sz = SIZES[args.workload]
    log(f"workload {args.workload}: {sz}")
    t0 = time.perf_counter()
    sys_ = make_system(sz["nocc"], sz["nvirt"], sz["naux"], seed=0)
    tg = time.perf_counter() - t0
    t0 = time.perf_counter()
    op = lib.Operator.from_system(sys_, keep_lowprec=True, coupling=args.coupling)
    t_build = time.perf_counter() - t0
    form = op.coupling
    log(f"system generated in {tg:.1f}s; operator built on GPU in {t_build:.2f}s (ndim {op.ndim}, "
        f"coupling form {form})")
    hip = Hip()
    nvec = args.nvec
    nbytes = op.ndim * nvec * 8
    dX, dW = hip.malloc(nbytes), hip.malloc(nbytes)
    X = np.random.default_rng(0).standard_normal((nvec, op.ndim))  # column-major ndim x nvec
    hip.h2d(dX, X)
    del X

    secs, prof = timed_sigma(op, hip, dX, dW, nvec, "fp64", args.steps, args.warmup)
    ms_per_step = secs / args.steps * 1e3
    value = nvec * args.steps / secs
    roof = roofline_of(prof, "fp64", args.workload, form)
    streams = stream_stats(prof)
    log(f"fp64: {ms_per_step:.3f} ms/step, {value:.1f} sigma-vectors/s, coupling {roof['achieved']} TFLOP/s, "
        f"minor {minor_stats(prof)}, streaming {streams}")
    extra = {"fp64": {"ms_per_step": round(ms_per_step, 4), "sigma_vectors_per_s": round(value, 2),
                      "coupling": roof, "coupling_minor": minor_stats(prof), "streaming": streams}}
    if not args.skip_lowprec:
        for prec in ("fp32", "bf16"):
            s2, p2 = timed_sigma(op, hip, dX, dW, nvec, prec, max(2, args.steps // 2), 1)
            r2 = roofline_of(p2, prec, args.workload, form)
            v2 = nvec * max(2, args.steps // 2) / s2
            extra[prec] = {"ms_per_step": round(s2 / max(2, args.steps // 2) * 1e3, 4),
                           "sigma_vectors_per_s": round(v2, 2), "coupling": r2,
                           "coupling_minor": minor_stats(p2), "streaming": stream_stats(p2)}
            log(f"{prec}: {extra[prec]['ms_per_step']} ms/step, {v2:.1f} sigma-vectors/s, coupling "
                f"{r2['achieved']} TFLOP/s")
    hip.free(dX)
    hip.free(dW)
    if args.copy_calibration:
        copy_calibration(hip)

    cpu = cpu_roots = None
    if not args.skip_cpu:
        try:
            cpu = cpu_baseline(sys_, op, nvec)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e}")
        try:
            cpu_roots = cpu_time_to_roots()
        except Exception as e:  # noqa: BLE001 - reported, never fatal
            log(f"cpu time-to-roots failed: {e}")
            cpu_roots = {"error": str(e)}
    dav = None
    if not args.skip_davidson:
        # the large workload's 40 roots: the block must hold every root
        # (block_davidson.f90:195, 221-222), and two blocks of 40 FP64
        # vectors of ndim 72,018,000 are what HBM holds beside the operator
        # (V and W 46 GB each, the Ritz, W y and residual blocks 23 GB each);
        # the diagonal steps do not converge there in 1000 iterations
        large = args.workload == "large"
        nroots = args.dav_roots or (40 if large else 20)
        block = args.dav_block or nroots
        maxvec = args.dav_maxvec or (2 * block if large else 10 * block)
        runs = args.dav_runs or ("fp64_coupled,fp32_coupled_refined" if large else "")
        dav = davidson_compare(op, nroots, [r for r in runs.split(",") if r], blocksize=block, maxvec=maxvec)

    sil_t = args.sil if args.sil >= 0 else (10.0 if args.workload == "benzene" else 0.0)
    sil = None
    if sil_t > 0:
        try:
            sil = sil_compare(op, sz, sil_t, args.sil_tol)
        except Exception as e:  # noqa: BLE001 - reported, never fatal
            log(f"sil failed: {e}")
            sil = {"error": str(e)}
    def guarded(name, fn):
        """Run one consumer measurement; a failure is reported in the JSON
        line instead of ending the benchmark."""
        try:
            return fn()
        except Exception as e:  # noqa: BLE001 - reported, never fatal
            log(f"{name} failed: {e}")
            return {"error": str(e)}

    cheb = guarded("cheby", lambda: cheby_measure(op, sz, args.cheby)) if args.cheby > 0 else None
    lanc = guarded("lanczos", lambda: lanczos_measure(op, sz, args.lanczos)) if args.lanczos else None
    rlx = guarded("relax", lambda: relax_measure(op, args.relax)) if args.relax > 0 else None
    mp2r = guarded("mp2", lambda: mp2_measure(op)) if not args.skip_mp2 else None
    fvr = guarded("fvector", lambda: fvector_measure(op)) if not args.skip_fvector else None
    las = guarded("laser", lambda: laser_measure(op, args.laser)) if args.laser > 0 else None
    tdr = guarded("trdens", lambda: trdens_measure(op)) if args.trdens else None
    fds = guarded("fdstates", lambda: fdstates_measure(op, sz, args.fdstates)) if args.fdstates > 0 else None
    # last: the ADC(2)-x leg switches the operator and back
    adx = guarded("adc2x", lambda: adc2x_measure(op, roots=max(args.adc2x_roots, 0))) \
        if args.adc2x_roots >= 0 else None

    n1 = op.n1
    bytes_per_sigma = 8.0 * (2 * op.ndim) + 8.0 * 2 * n1 * n1  # read x, write sigma, pair tensors
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "sigma-vectors/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded orbital energies and density-fitting factors, seed 0)",
        "config": {"workload": f"{args.workload}: ADC(2)-s singlet o={sz['nocc']} v={sz['nvirt']} "
                               f"naux={sz['naux']}, ndim={op.ndim}, block of {nvec} vectors per step",
                   "nvec": nvec, "ndim": op.ndim, "precision": "fp64", "coupling_form": form},
        "roofline": {k: roof[k] for k in ("bound", "achieved", "peak", "unit", "frac", "traffic")},
        "cpu_baseline": cpu,
        "sigma_GBps_effective": round(bytes_per_sigma * value / 1e9, 1),
        "detail": extra,
        "davidson_20_roots": dav,
        "cpu_time_to_5_roots_water": cpu_roots,
        "sil_autocorrelation": sil,
        "chebyshev_autocorrelation": cheb,
        "block_lanczos": lanc,
        "relaxation": rlx,
        "mp2": mp2r,
        "fvector": fvr,
        "laser_propagation": las,
        "transition_density": tdr,
        "cheby_fdstates": fds,
        "adc2x": adx,
        "operator_build_seconds": round(t_build, 3),
    }
    op.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
