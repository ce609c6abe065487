"""Seeded synthetic closed-shell systems (orbital energies + density-fitted factors).

The reference reads Hartree-Fock orbitals and the full nbas^4 MO integral array
from GAMESS (source/qclib/load_electronic_structure.f90:153-342,
source/multi/integrals_mo2e.f90:107-122).  This module replaces that reader by
a seeded generator of the quantities the ADC(2) sigma-vector needs:

* ``eps``: orbital energies, occupied first (ascending, U[-1.5,-0.3] Eh), then
  virtual (ascending, U[0.05,3.0] Eh);
* ``B[P,p,q]``: density-fitting factors, N(0, 0.05^2), symmetric in (p,q), so
  that (pq|rs) = sum_P B[P,p,q] B[P,r,s] carries the real 8-fold symmetry the
  reference's ``vpqrs`` lookup assumes (source/qclib/vpqrsmod.f90:13-27).

Draw order (fixed, so every consumer builds the same system from a seed):
numpy ``Generator(PCG64(seed))``; occupied energies, virtual energies, then the
upper triangle (p<=q, row-major) of B for P = 0..naux-1.
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class System:
    nocc: int
    nvirt: int
    naux: int
    eps: np.ndarray  # (nbas,) float64
    B: np.ndarray    # (naux, nbas, nbas) float64, C-contiguous

    @property
    def nbas(self) -> int:
        return self.nocc + self.nvirt


# Named sizes of the workloads this package is measured on.
SIZES = {
    "water": dict(nocc=5, nvirt=36, naux=164),
    "benzene": dict(nocc=16, nvirt=200, naux=864),
    "large": dict(nocc=30, nvirt=400, naux=1720),
}


def make_system(nocc: int, nvirt: int, naux: int, seed: int = 0) -> System:
    if nocc < 2 or nvirt < 2 or naux < 1:
        raise ValueError("need nocc>=2, nvirt>=2, naux>=1")
    rng = np.random.Generator(np.random.PCG64(seed))
    eo = np.sort(rng.uniform(-1.5, -0.3, nocc))
    ev = np.sort(rng.uniform(0.05, 3.0, nvirt))
    nbas = nocc + nvirt
    iu, ju = np.triu_indices(nbas)
    B = np.empty((naux, nbas, nbas), dtype=np.float64)
    # draw per P block to bound peak memory at large sizes
    for P in range(naux):
        vals = rng.normal(0.0, 0.05, size=iu.size)
        B[P, iu, ju] = vals
        B[P, ju, iu] = vals
    return System(nocc, nvirt, naux, np.concatenate([eo, ev]), B)


def unit_dipole(nocc: int, nvirt: int, seed: int = 0) -> np.ndarray:
    """Synthetic transition-dipole amplitudes mu[a,i] ~ N(0, 0.1^2) in the
    1h1p ordering (virtual outer, occupied inner), used as the wavepacket
    initial state (the zeroth-order F-vector term of
    source/adclib/get_moment.f90:186)."""
    rng = np.random.Generator(np.random.PCG64(seed + 7919))
    return rng.normal(0.0, 0.1, size=nocc * nvirt)
