"""ADC(2)-s (or, with --adc2x, ADC(2)-x) excitation energies of a molecule
from its GAMESS output.

    python -m adc_ottawa_amd.frontend gamess.log --frozen 11 --roots 10 [--cvs] [--adc2x] [--block 20] [--maxvec 60]
    python -m adc_ottawa_amd.frontend gamess.log --roots 5 --nto

Reads the basis and geometry from the GAMESS log, computes the AO integrals,
redoes the RHF, factorises the MO integrals (pivoted Cholesky to 1e-11 on the diagonal) and
runs on the GPU library: the frozen-core MP2 energy, then the block Davidson
(Olsen step and ADC(1) guess, the reference's defaults) on the frozen-core
view of the configurations -- the reference's initial space with frozen_core
-- or, with --cvs, on the CVS view with the frozen orbitals as core orbitals
-- the reference's final space with cvs.  Energies are printed in eV
(1 Eh = 27.2113845 eV, the reference's constant).  With --nto each state's
natural transition orbitals as adc2_nto_gs forms them: the
pairs with sigma^2 >= 0.01, hole and particle orbitals in the AO basis.
"""
from __future__ import annotations

import argparse
import json
import sys

from .adc import run


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("log")
    ap.add_argument("--frozen", type=int, default=0, help="frozen (core) orbitals, the lowest occupied")
    ap.add_argument("--roots", type=int, default=10)
    ap.add_argument("--cvs", action="store_true", help="CVS space on the frozen orbitals")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--maxvec", type=int, default=0)
    ap.add_argument("--tol", type=float, default=1e-7)
    ap.add_argument("--precon", default="olsen", choices=("dpr", "olsen"))
    ap.add_argument("--guess", default="adc1", choices=("adc1", "ondiag", "subdiag"))
    ap.add_argument("--deflate", action="store_true")
    ap.add_argument("--adc2x", action="store_true", help="ADC(2)-x: the first-order 2h2p-2h2p block")
    ap.add_argument("--nto", action="store_true",
                    help="natural transition orbitals of each state: sigma^2 and AO coefficients")
    a = ap.parse_args()
    r = run(a.log, a.frozen, a.roots, a.cvs, a.block, a.maxvec, a.tol, a.precon, a.guess, a.deflate, adc2x=a.adc2x,
            nto=a.nto)
    print(json.dumps(r, indent=1))
    print("excitation energies (eV): " + " ".join(f"{e:.6f}" for e in r["energies_eV"]), file=sys.stderr)
    return 0


if __name__ == "__main__":
    sys.exit(main())
