"""Configuration (CSF) tables of the ADC(2)-s singlet space.  Test oracle only.

Restates the ordering of source/adclib/select_fano.f90 for a closed shell with
C1 symmetry (nirrep=1, hcentre = all occupied, roccnum = identity):

* 1h1p  (select_atom_is, select_fano.f90:16-60): virtual ``a`` outer,
  occupied ``i`` inner.
* 2h2p  (select_atom_d flag=-1, select_fano.f90:539-718):
  t1  a=b, i=j   : i outer, a inner                         (:569-592)
  t2  a<b, i=j   : a outer, b>a, i inner                    (:597-625)
  t3  a=b, i<j   : i outer, j>i, a inner                    (:628-655)
  t4  a<b, i<j I : i outer, j>i, a, b>a (spin 11)           (:658-696)
  t5  a<b, i<j II: a copy of the t4 block with type 5       (:700-702)

Indices here are 0-based: ``a, b`` in [0, nvirt), ``k, l`` in [0, nocc).  The
reference column (cnf, spin, j, k, a, b, type) of ``fill_indices``
(source/utilities/misc.f90:458-471) maps to (k, l, a, b) = (col3, col4, col5,
col6) minus the 1-based/occupied offsets.

Pinned by: the per-type counts of the survey's size table, and by
``test_oracle_csf`` (bijectivity and the closed-form index maps).
"""
from __future__ import annotations

import numpy as np


def counts(o: int, v: int) -> dict:
    npv = v * (v - 1) // 2
    npo = o * (o - 1) // 2
    c = {"n1": o * v, "t1": o * v, "t2": npv * o, "t3": npo * v,
         "t4": npo * npv, "t5": npo * npv}
    c["n2"] = c["t1"] + c["t2"] + c["t3"] + c["t4"] + c["t5"]
    c["ndim"] = c["n1"] + c["n2"]
    return c


def table(o: int, v: int):
    """Return (typ, a, b, k, l) int arrays of length ndim in reference order.

    typ: 0 for 1h1p, 1..5 for the 2h2p types.  For 1h1p, b = l = -1.
    """
    rows = []
    # 1h1p: a outer, i inner (select_fano.f90:45-56)
    for a in range(v):
        for i in range(o):
            rows.append((0, a, -1, i, -1))
    # t1: i outer, a inner
    for i in range(o):
        for a in range(v):
            rows.append((1, a, a, i, i))
    # t2: a outer, b>a, i inner
    for a in range(v):
        for b in range(a + 1, v):
            for i in range(o):
                rows.append((2, a, b, i, i))
    # t3: i outer, j>i, a inner
    for i in range(o):
        for j in range(i + 1, o):
            for a in range(v):
                rows.append((3, a, a, i, j))
    # t4: i, j>i, a, b>a
    t4 = []
    for i in range(o):
        for j in range(i + 1, o):
            for a in range(v):
                for b in range(a + 1, v):
                    t4.append((4, a, b, i, j))
    rows.extend(t4)
    # t5: copy of t4 with type 5 (select_fano.f90:700-702)
    rows.extend((5,) + r[1:] for r in t4)
    arr = np.array(rows, dtype=np.int64)
    return arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], arr[:, 4]


def cvs_indices(o: int, v: int, core) -> np.ndarray:
    """Full-space indices of the CVS configurations, in the CVS order.

    The loops follow select_atom_is_cvs (source/adclib/select_fano.f90:112-142)
    and select_atom_d_cvs (:903-1027).  1h1p: a outer, core hole i inner.  No
    t1 or t2.  t3: pairs i < j with exactly one core orbital, a inner.  t4,
    then t5: the same pairs, (a, b) inner.  ``core`` holds 0-based occupied
    indices.
    """
    c = counts(o, v)
    isc = np.zeros(o, dtype=int)
    isc[np.asarray(core, dtype=int)] = 1
    npv = v * (v - 1) // 2
    off2 = c["n1"] + c["t1"] + c["t2"]
    off3 = off2 + c["t3"]
    off4 = off3 + c["t4"]
    out = [a * o + i for a in range(v) for i in range(o) if isc[i]]
    pairs = [(i, j) for i in range(o) for j in range(i + 1, o)]
    for p, (i, j) in enumerate(pairs):
        if isc[i] + isc[j] == 1:
            out.extend(off2 + p * v + a for a in range(v))
    for off in (off3, off4):
        for p, (i, j) in enumerate(pairs):
            if isc[i] + isc[j] == 1:
                out.extend(off + p * npv + q for q in range(npv))
    return np.asarray(out, dtype=np.int64)


def frozen_core_indices(o: int, v: int, frozen) -> np.ndarray:
    """Full-space indices of the frozen-core configurations, in the full
    order: those none of whose holes is a frozen orbital (the lifrzcore
    cycles of select_atom_is / select_atom_d, source/adclib/select_fano.f90:
    27-40, 575, 610, 635-642, 665-672).  ``frozen`` holds 0-based occupied
    indices."""
    typ, a, b, k, l = table(o, v)
    isf = np.zeros(o, dtype=bool)
    isf[np.asarray(frozen, dtype=int)] = True
    bad = isf[k] | ((l >= 0) & isf[np.maximum(l, 0)])
    return np.where(~bad)[0].astype(np.int64)
