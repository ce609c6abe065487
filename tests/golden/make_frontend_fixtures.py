"""Write the molecular-front-end fixtures from the reference's examples.

    python tests/golden/make_frontend_fixtures.py

Reads (as text) examples/cs2_sto2g and examples/n2_augccpvdz under
/root/reference and writes, per example (for examples/h2o_autospec and
examples/allene_lanczos, which ship no ADC output, only the first file, as
h2o_augccpvdz_gamess.txt and allene_ccpvdz_gamess.txt):
* <name>_gamess.txt: the GAMESS log sections the front end reads (the
  number of electrons, the 'COORDINATES (BOHR)' table, the 'ATOMIC BASIS SET'
  block) and GAMESS's final RHF and nuclear-repulsion energies -- input data,
  copied line for line;
* <name>_reference.json: numbers the reference printed in its own run on
  that input (diff.*.log): HF and MP2 energies, the hole/particle space sizes,
  the configuration-space dimensions, and the converged block-Davidson
  energies (eV) of the initial (valence, frozen core) and final (CVS) spaces,
  each root taken at the iteration where it is marked converged, and the
  Davidson settings of the reference's input (diag_section for the initial
  space, diag_final_section for the final space).
"""
from __future__ import annotations

import json
import os
import re

REF = "/root/reference/examples"
OUT = os.path.dirname(os.path.abspath(__file__))
EXAMPLES = {"cs2_sto2g": ("diff.cs2.log", "cs2_adc2.inp"), "n2_augccpvdz": ("diff.adc_z.log", "adc_z.inp")}
# examples that ship only the GAMESS run (no ADC output): the log sections only
GAMESS_ONLY = {"h2o_augccpvdz": "h2o_autospec", "allene_ccpvdz": "allene_lanczos"}


def trimmed_log(path):
    lines = open(path).read().splitlines()
    keep = []
    i = next(k for k, s in enumerate(lines) if "COORDINATES (BOHR)" in s and "ATOM" in s)
    j = i + 2
    while lines[j].strip():
        j += 1
    keep += lines[i:j + 1]
    i = next(k for k, s in enumerate(lines) if s.strip() == "ATOMIC BASIS SET")
    j = next(k for k in range(i, len(lines)) if "TOTAL NUMBER OF BASIS SET SHELLS" in lines[k])
    keep += lines[i:j + 1]
    for pat in ("NUMBER OF CARTESIAN GAUSSIAN BASIS FUNCTIONS", "NUMBER OF ELECTRONS",
                "THE NUCLEAR REPULSION ENERGY IS", "FINAL RHF ENERGY IS"):
        keep.append(next(s for s in lines if pat in s))
    return "\n".join(keep) + "\n"


def davidson_blocks(lines):
    """Converged energies (eV) of each 'Block Davidson diagonalisation in the
    <space> space' table, in the order they converged."""
    out = {}
    for k, s in enumerate(lines):
        m = re.search(r"Block Davidson diagonalisation in the (\w+) space", s)
        if not m:
            continue
        conv = []
        for t in lines[k + 1:]:
            if "All roots converged" in t:
                break
            f = t.split()
            if len(f) >= 3 and f[-1] == "y":
                conv.append(float(f[-3]))
        out[m.group(1)] = conv
    return out


def reference_numbers(path):
    lines = open(path).read().splitlines()
    text = "\n".join(lines)
    num = lambda pat: float(re.search(pat, text).group(1))  # noqa: E731
    dims = {}
    for key in ("INITIAL", "FINAL", "TOTAL"):
        k = next(i for i, s in enumerate(lines) if f"dimension of various {key} configuration spaces" in s)
        dims[key.lower()] = [int(x) for x in lines[k + 1].split()]
    return {
        "hf_energy": num(r"\* HF energy:\s+(-?\d+\.\d+)"),
        "mp2_correlation": num(r"\* MP2 correlation energy:\s+(-?\d+\.\d+)"),
        "hole_space": int(num(r"hole space size\s+(\d+)")),
        "particle_space": int(num(r"particle space size\s+(\d+)")),
        "configuration_counts": dims,
        "davidson_eV": davidson_blocks(lines),
        "source": path.replace("/root/reference/", ""),
    }


def davidson_settings(path):
    """{space: {nstates, block_size, maxvec, tol, preconditioner, guess, deflate}}."""
    out, cur = {}, None
    for s in open(path).read().splitlines():
        t = s.split("#")[0].strip()
        if t in ("diag_section", "diag_final_section"):
            cur = {"deflate": False}
            out["initial" if t == "diag_section" else "final"] = cur
        elif t.startswith("end-diag"):
            cur = None
        elif cur is not None and t:
            if t == "deflate":
                cur["deflate"] = True
            elif "=" in t:
                k, v = (x.strip() for x in t.split("=", 1))
                cur[k] = float(v) if k == "tol" else (int(v) if v.isdigit() else v)
    return out


def main():
    for name, (diff, inp) in EXAMPLES.items():
        log = trimmed_log(os.path.join(REF, name, "gamess.log"))
        with open(os.path.join(OUT, f"{name}_gamess.txt"), "w") as f:
            f.write(log)
        ref = reference_numbers(os.path.join(REF, name, diff))
        ref["davidson_settings"] = davidson_settings(os.path.join(REF, name, inp))
        with open(os.path.join(OUT, f"{name}_reference.json"), "w") as f:
            json.dump(ref, f, indent=1)
        print(name, {k: (v if k != "davidson_eV" else {s: len(e) for s, e in v.items()}) for k, v in ref.items()})
    for name, src in GAMESS_ONLY.items():
        with open(os.path.join(OUT, f"{name}_gamess.txt"), "w") as f:
            f.write(trimmed_log(os.path.join(REF, src, "gamess.log")))
        print(name, "GAMESS log sections")


if __name__ == "__main__":
    main()
