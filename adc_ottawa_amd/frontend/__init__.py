"""Molecular front end: GAMESS basis and geometry -> AO integrals -> RHF ->
MO integrals -> density-fitting factors B[P,p,q] for the ADC(2) library.

Replaces the reference's `load_gamess` / `load_mo_integrals`
(source/iomodules/read_param.f90:14-61,
source/qclib/load_electronic_structure.f90:153-342), which compute the AO
integrals of the GAMESS basis with the reference's own integral code, redo
the RHF, and keep the full nbas^4 MO integral array
(source/multi/integrals_mo2e.f90:107-122).  Here the MO integrals are
factorised exactly (pivoted Cholesky to a set tolerance) into the B[P,p,q]
the library takes.  Host code: numpy plus a C integral engine (ints.c),
built with the package.
"""
from .gamess import Molecule, read_gamess_log  # noqa: F401
from .scf import rhf, mo_system  # noqa: F401
