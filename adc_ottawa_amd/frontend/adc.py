"""ADC(2)-s or ADC(2)-x of a molecule through the GPU library: the
frozen-core MP2 energy and the block-Davidson roots of the reference's
initial (frozen-core) or final (CVS) configuration space, with the
reference's ADC(1) guess.  ADC(2)-x (the reference's method = adc2-x) adds
the first-order 2h2p-2h2p block (adc2g_set_adc2x) before the view is made.

The reference's guess = adc1 (initvec_adc1, source/eigen/block_davidson.f90:
246-289) starts the Davidson from the lowest eigenvectors of the ADC(1)
matrix of the space's 1h1p configurations (zero on the 2h2p ones); here that
matrix is the library's first-order 1h1p block (adc2g_phph of a handle made
without the second-order terms) restricted to those configurations.
"""
from __future__ import annotations

import time

import numpy as np

from .. import lib
from . import gamess, scf

EH2EV = 27.2113845


def view_1h1p_rows(nocc: int, nvirt: int, kind: str, orbitals) -> np.ndarray:
    """Full-space 1h1p indices (a*nocc + i) of a view's 1h1p configurations,
    in the view's order: a outer, i inner; kind 'frozen' keeps the holes not in
    ``orbitals``, 'cvs' the holes in them (1-based occupied indices)."""
    sel = np.zeros(nocc, dtype=bool)
    sel[np.asarray(orbitals, dtype=int) - 1] = True
    keep = ~sel if kind == "frozen" else sel
    return np.array([a * nocc + i for a in range(nvirt) for i in range(nocc) if keep[i]], dtype=np.int64)


def adc1_guess(system, kind: str, orbitals, ndim: int, nblock: int) -> np.ndarray:
    """(ndim, nblock) orthonormal guess: the nblock lowest ADC(1) eigenvectors
    of the view's 1h1p configurations (the first rows of the view), zero on
    its 2h2p configurations."""
    with lib.Operator.from_system(system, second_order=False) as op1:
        A = op1.phph()
    rows = view_1h1p_rows(system.nocc, system.nvirt, kind, orbitals) if kind else np.arange(A.shape[0])
    if nblock > rows.size:
        raise ValueError(f"the ADC(1) guess has {rows.size} vectors, fewer than the block size {nblock}")
    w, U = np.linalg.eigh(A[np.ix_(rows, rows)])
    G = np.zeros((ndim, nblock))
    G[: rows.size] = U[:, :nblock]
    return G


def run(log, frozen, roots, cvs=False, block=0, maxvec=0, tol=1e-7, precon="olsen", guess="adc1", deflate=False,
        verbose=False, adc2x=False, nto=False):
    t0 = time.perf_counter()
    mol = gamess.read_gamess_log(log)
    ints = gamess.ao_integrals(mol)
    hf = scf.rhf(mol, ints)
    full, res_full = scf.mo_system(hf, ints, frozen=0)
    out = {"nao": mol.nao, "nocc": hf.nocc, "hf_energy": hf.energy, "scf_iterations": hf.iterations,
           "naux": full.naux, "cholesky_max_error": res_full, "frozen": frozen}
    act = scf.mo_system(hf, ints, frozen=frozen)[0] if frozen > 0 else full
    with lib.Operator.from_system(act) as op:
        e2, d2, _ = op.mp2(density=False)
    out["mp2_correlation"], out["d2"] = e2, d2
    orbs = list(range(1, frozen + 1))
    kind = "cvs" if cvs else ("frozen" if frozen > 0 else None)
    block = block or roots
    maxvec = maxvec or 4 * block
    with lib.Operator.from_system(full) as op:
        if adc2x:
            op.set_adc2x(True)
        out["method"] = "adc2-x" if adc2x else "adc2-s"
        view = op.cvs(orbs) if kind == "cvs" else (op.frozen_core(orbs) if kind == "frozen" else None)
        target = view if view is not None else op
        try:
            g = adc1_guess(full, kind, orbs, target.ndim, block) if guess == "adc1" else guess
            ev, vecs, st = target.davidson(roots, blocksize=block, maxvec=maxvec, tol=tol, precon=precon,
                                           deflate=deflate, maxiter=1000, vectors=nto, verbose=verbose, guess=g)
            if nto:
                # adc2_nto_gs (nto.f90:76-260): the SVD of each state's
                # occupied-virtual transition density, pairs with sigma^2 >= 0.01;
                # hole and particle orbitals in the AO basis
                T = target.trdens_vo(vecs)
                out["nto"] = []
                for s_ in range(T.shape[0]):
                    sig, hole, part = lib.nto(T[s_], hf.C[:, :hf.nocc], hf.C[:, hf.nocc:], thresh=0.01)
                    out["nto"].append({"sigma2": [float(x) ** 2 for x in sig[: hole.shape[1]]],
                                       "hole_ao": hole.tolist(), "particle_ao": part.tolist()})
            out["space"] = {"cvs": "cvs", "frozen": "frozen core", None: "full"}[kind]
            out["ndim"] = target.ndim
            out["counts"] = [target.n1] + list(target.counts)
            out["energies_eV"] = [float(x) * EH2EV for x in ev]
            out["davidson_iterations"] = st["niter"]
        finally:
            if view is not None:
                view.close()
    out["seconds"] = time.perf_counter() - t0
    return out
