"""Molecular front end on the CPU: the AO integral engine against closed
forms and rotational invariance, the RHF against GAMESS's energies for the
reference's examples, the Cholesky factors against the MO integrals, and
the frozen-core MP2 energy and valence ADC(2)-s roots of CS2 (the oracle's
matrix on the frozen-core configurations) against the numbers the
reference printed for the same input (tests/golden, written by
tests/golden/make_frontend_fixtures.py)."""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sl
import scipy.special as sp

from adc_ottawa_amd.frontend import gamess, scf
from oracle import csf, elements, mp2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
EH2EV = 27.2113845


def _mol(centres, shells, nelec=2):
    """Molecule from [(x, y, z, charge)] and [(atom, l, exps, coefs)]."""
    xyz = np.array([c[:3] for c in centres], dtype=float)
    Z = np.array([c[3] for c in centres], dtype=float)
    sh = [gamess.Shell(at, l, np.array(e, float), np.array(c, float)) for at, l, e, c in shells]
    return gamess.Molecule(["X"] * len(centres), Z, xyz, sh, nelec)


def _boys0(t):
    return 1.0 if t < 1e-14 else 0.5 * np.sqrt(np.pi / t) * sp.erf(np.sqrt(t))


def test_s_functions_closed_forms():
    """Normalised s primitives a at A, b at B and a unit charge at C: overlap,
    kinetic energy, nuclear attraction and (aa|bb) by their closed forms."""
    a, b = 0.8, 1.7
    A, B, Cc = np.array([0.0, 0.0, 0.0]), np.array([0.3, -0.5, 1.1]), np.array([-0.7, 0.4, 0.2])
    m = _mol([(*A, 0.0), (*B, 0.0), (*Cc, 1.0)], [(0, 0, [a], [1.0]), (1, 0, [b], [1.0])])
    ints = gamess.ao_integrals(m)
    p, mu, R2 = a + b, a * b / (a + b), np.sum((A - B) ** 2)
    S = (2.0 * np.sqrt(a * b) / p) ** 1.5 * np.exp(-mu * R2)
    T = mu * (3.0 - 2.0 * mu * R2) * S
    P = (a * A + b * B) / p
    V = -2.0 * np.sqrt(p / np.pi) * S * _boys0(p * np.sum((P - Cc) ** 2))
    assert abs(ints["S"][0, 1] - S) < 1e-14 and abs(ints["S"][0, 0] - 1.0) < 1e-14
    assert abs(ints["T"][0, 1] - T) < 1e-13
    assert abs(ints["V"][0, 1] - V) < 1e-13
    # (aa|bb): charge distributions of exponents 2a at A and 2b at B
    pa, pb = 2 * a, 2 * b
    eri = 2.0 * np.sqrt(pa * pb / (np.pi * (pa + pb))) * _boys0(pa * pb / (pa + pb) * R2)
    assert abs(ints["eri"][0, 0, 1, 1] - eri) < 1e-13
    # dipole of the overlap distribution: S * P
    assert np.abs(ints["dipole"][:, 0, 1] - S * P).max() < 1e-14


def test_integrals_rotation_invariant():
    """The RHF energy of a molecule with s, p and d shells on three centres
    is unchanged by a rotation of the frame (1e-10)."""
    rng = np.random.default_rng(3)
    cen = [(0.0, 0.0, 0.0, 3.0), (1.4, 0.3, -0.2, 1.0), (-0.6, 1.1, 0.5, 2.0)]
    shells = [(0, 0, [4.0, 0.9], [0.5, 0.6]), (0, 1, [1.1], [1.0]), (0, 2, [0.8], [1.0]),
              (1, 0, [1.3], [1.0]), (1, 1, [0.7, 2.0], [0.7, 0.4]), (2, 0, [2.2, 0.5], [0.4, 0.7]),
              (2, 2, [1.2], [1.0])]
    Q = np.linalg.qr(rng.standard_normal((3, 3)))[0]
    energies = []
    for R in (np.eye(3), Q):
        m = _mol([tuple(R @ np.array(c[:3])) + (c[3],) for c in cen], shells, nelec=6)
        energies.append(scf.rhf(m, gamess.ao_integrals(m)).energy)
    assert abs(energies[0] - energies[1]) < 1e-10


@pytest.mark.parametrize("name", ["cs2_sto2g", "n2_augccpvdz", "h2o_augccpvdz", "allene_ccpvdz"])
def test_rhf_matches_gamess(name):
    log = os.path.join(GOLDEN, f"{name}_gamess.txt")
    lines = open(log).read().splitlines()
    e_gamess = float(next(s for s in lines if "FINAL RHF ENERGY IS" in s).split()[4])
    e_nuc = float(next(s for s in lines if "NUCLEAR REPULSION ENERGY IS" in s).split()[-1])
    mol = gamess.read_gamess_log(log)
    assert mol.nao == int(next(s for s in lines if "CARTESIAN GAUSSIAN BASIS FUNCTIONS" in s).split()[-1])
    assert abs(mol.nuclear_repulsion() - e_nuc) < 1e-9
    ints = gamess.ao_integrals(mol)
    hf = scf.rhf(mol, ints)
    assert abs(hf.energy - e_gamess) < 1e-7


def test_cs2_mp2_and_valence_adc2_vs_reference():
    """CS2/STO-2G+: frozen-core MP2 correlation energy and the 10 lowest
    valence ADC(2)-s roots of the frozen-core space (the oracle's matrix of
    the whole system on those configurations) against the reference's run."""
    log = os.path.join(GOLDEN, "cs2_sto2g_gamess.txt")
    ref = json.load(open(os.path.join(GOLDEN, "cs2_sto2g_reference.json")))
    mol = gamess.read_gamess_log(log)
    ints = gamess.ao_integrals(mol)
    hf = scf.rhf(mol, ints)
    nfrozen = ref["hole_space"] - ref["configuration_counts"]["initial"][0] // ref["particle_space"]
    assert nfrozen == 11
    full, res = scf.mo_system(hf, ints, frozen=0)
    assert res < 1e-9
    act, _ = scf.mo_system(hf, ints, frozen=nfrozen)
    e2 = mp2.energy(act, mp2.integrals(act))
    assert abs(e2 - ref["mp2_correlation"]) < 2e-7
    o, v = full.nocc, full.nvirt
    idx = csf.frozen_core_indices(o, v, range(nfrozen))
    typ = csf.table(o, v)[0][idx]
    counts = [int(np.sum(typ == t)) for t in range(6)]
    assert counts == ref["configuration_counts"]["initial"] + ref["configuration_counts"]["initial"][-1:]
    blk = elements.build_blocks(full, apply_minc=False)
    n1 = blk.n1
    i1, i2 = idx[idx < n1], idx[idx >= n1] - n1
    H = np.block([[blk.H11[np.ix_(i1, i1)], blk.C[np.ix_(i1, i2)]],
                  [blk.C[np.ix_(i1, i2)].T, np.diag(blk.D[i2])]])
    w = sl.eigh(H, eigvals_only=True, subset_by_index=[0, 9]) * EH2EV
    want = np.sort(ref["davidson_eV"]["initial"])
    assert np.abs(w - want).max() < 2e-5
