"""GAMESS output reader and AO integrals.

The geometry is read from the log's 'COORDINATES (BOHR)' table (GAMESS's own
Angstrom-to-bohr conversion) and the basis from its 'ATOMIC BASIS SET' block
(contraction coefficients of normalised primitives, 12 digits), one block
per atom in input order -- the data the reference's `import_gamess` takes
from the same files.  Cartesian AOs (ISPHER=-1) in GAMESS order: s; x y z;
xx yy zz xy xz yz.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libadc2g_ints.so")
LVAL = {"S": 0, "P": 1, "D": 2}


@dataclasses.dataclass
class Shell:
    atom: int
    l: int
    exps: np.ndarray
    coefs: np.ndarray


@dataclasses.dataclass
class Molecule:
    symbols: list
    charges: np.ndarray  # (natom,)
    xyz: np.ndarray      # (natom, 3) bohr
    shells: list
    nelec: int

    @property
    def nao(self) -> int:
        return sum((s.l + 1) * (s.l + 2) // 2 for s in self.shells)

    def nuclear_repulsion(self) -> float:
        e = 0.0
        for i in range(len(self.charges)):
            for j in range(i):
                e += self.charges[i] * self.charges[j] / np.linalg.norm(self.xyz[i] - self.xyz[j])
        return float(e)


def read_gamess_log(path: str) -> Molecule:
    with open(path) as f:
        lines = f.read().splitlines()
    # geometry (bohr)
    i = next(k for k, s in enumerate(lines) if "COORDINATES (BOHR)" in s and "ATOM" in s)
    symbols, charges, xyz = [], [], []
    k = i + 2
    while lines[k].strip():
        t = lines[k].split()
        symbols.append(t[0])
        charges.append(float(t[1]))
        xyz.append([float(x) for x in t[2:5]])
        k += 1
    natom = len(symbols)
    # basis: one block per atom, shells numbered, primitives one per line
    i = next(k for k, s in enumerate(lines) if s.strip() == "ATOMIC BASIS SET")
    k = next(j for j in range(i, len(lines)) if "SHELL TYPE" in lines[j]) + 1
    shells, atom, cur = [], -1, None
    while True:
        s = lines[k].strip()
        if s.startswith("TOTAL NUMBER OF BASIS SET SHELLS"):
            break
        t = s.split()
        if len(t) == 1 and t[0].isalpha():
            atom += 1
            cur = None
        elif len(t) >= 5 and t[0].isdigit():
            nsh, typ = int(t[0]), t[1]
            if typ not in LVAL:
                raise ValueError(f"shell type {typ} not supported (s, p, d)")
            if cur is None or cur[0] != nsh:
                cur = (nsh, Shell(atom, LVAL[typ], [], []))
                shells.append(cur[1])
            cur[1].exps.append(float(t[3]))
            cur[1].coefs.append(float(t[4]))
        k += 1
    if atom + 1 != natom:
        raise ValueError(f"{atom + 1} basis blocks for {natom} atoms")
    for sh in shells:
        sh.exps = np.array(sh.exps)
        sh.coefs = np.array(sh.coefs)
    nel = next(int(s.split("=")[1]) for s in lines if "NUMBER OF ELECTRONS" in s)
    return Molecule(symbols, np.array(charges), np.array(xyz), shells, nel)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built; run python -m adc_ottawa_amd.build")
        _lib = C.CDLL(LIB)
        for fn in (_lib.ints_one_electron, _lib.ints_eri):
            fn.restype = C.c_int
    return _lib


def _arrays(mol: Molecule):
    l = np.array([s.l for s in mol.shells], dtype=np.int32)
    atom = np.array([s.atom for s in mol.shells], dtype=np.int32)
    nprim = np.array([s.exps.size for s in mol.shells], dtype=np.int32)
    off = np.concatenate([[0], np.cumsum(nprim)[:-1]]).astype(np.int32)
    ao = np.concatenate([[0], np.cumsum((l + 1) * (l + 2) // 2)[:-1]]).astype(np.int32)
    exps = np.concatenate([s.exps for s in mol.shells]).astype(np.float64)
    coefs = np.concatenate([s.coefs for s in mol.shells]).astype(np.float64)
    xyz = np.ascontiguousarray(mol.xyz, dtype=np.float64)
    return l, atom, nprim, off, ao, exps, coefs, xyz


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def ao_integrals(mol: Molecule, eri: bool = True):
    """dict of S, T, V, dipole (3, nao, nao) and (if eri) the nao^4 repulsion
    integrals, every contracted Cartesian AO normalised to unity."""
    L = _load()
    l, atom, nprim, off, ao, exps, coefs, xyz = _arrays(mol)
    n = mol.nao
    i32, f64 = C.c_int, C.c_double
    S, T, V = (np.zeros((n, n)) for _ in range(3))
    D = np.zeros((3, n, n))
    Z = np.ascontiguousarray(mol.charges, dtype=np.float64)
    args = (len(mol.shells), _p(atom, i32), _p(l, i32), _p(nprim, i32), _p(off, i32), _p(ao, i32), _p(exps, f64),
            _p(coefs, f64), _p(xyz, f64))
    rc = L.ints_one_electron(*args, len(Z), _p(Z, f64), n, _p(S, f64), _p(T, f64), _p(V, f64), _p(D[0], f64),
                             _p(D[1], f64), _p(D[2], f64))
    if rc != 0:
        raise RuntimeError("one-electron integrals failed")
    s = 1.0 / np.sqrt(np.diag(S).copy())
    out = {"S": S * np.outer(s, s), "T": T * np.outer(s, s), "V": V * np.outer(s, s),
           "dipole": D * np.outer(s, s)[None]}
    if eri:
        G = np.zeros((n, n, n, n))
        rc = L.ints_eri(*args, n, _p(G, f64))
        if rc != 0:
            raise RuntimeError("repulsion integrals failed")
        out["eri"] = G * np.einsum("i,j,k,l->ijkl", s, s, s, s, optimize=True)
    return out
