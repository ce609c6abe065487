"""Build libadc2g.so (gfx950) in-tree with hipcc.

    python -m adc_ottawa_amd.build [--jobs N]

Compiles every ``csrc/*.hip`` to an object under ``adc_ottawa_amd/_build/``
(in parallel) and links ``adc_ottawa_amd/libadc2g.so``.  Packed-FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libadc2g.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src).replace(".hip", ".o"))
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    # rocSOLVER: the dense eigen-decomposition of the coupled preconditioner's Schur complement
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB




FORTRAN_DIR = os.path.join(HERE, "fortran")
DRIVER = os.path.join(HERE, "adc2g_driver")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = True) -> str:
    """Compile the iso_c_binding module and the Fortran driver against the
    in-tree libadc2g.so (rpath to the package directory)."""
    os.makedirs(BUILD, exist_ok=True)
    mod = os.path.join(FORTRAN_DIR, "adc2g_mod.f90")
    drv = os.path.join(FORTRAN_DIR, "adc2g_driver.f90")
    cmd = [FC, "-O2", "-J", BUILD, "-o", DRIVER, mod, drv, "-L" + HERE, "-ladc2g",
           "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=BUILD)
    if r.returncode != 0:
        raise RuntimeError(f"Fortran build failed:\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {DRIVER}", file=sys.stderr)
    return DRIVER


FRONTEND_DIR = os.path.join(HERE, "frontend")
INTS_LIB = os.path.join(FRONTEND_DIR, "libadc2g_ints.so")


def build_frontend(verbose: bool = True) -> str:
    """Compile the molecular front end's AO integral engine (host C, OpenMP)."""
    src = os.path.join(FRONTEND_DIR, "ints.c")
    tmp = INTS_LIB + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O3", "-fPIC", "-shared", "-fopenmp", "-Wall", "-o", tmp, src, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"front-end build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, INTS_LIB)
    if verbose:
        print(f"built {INTS_LIB}", file=sys.stderr)
    return INTS_LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--no-fortran", action="store_true")
    a = ap.parse_args()
    build(a.jobs)
    build_frontend()
    if not a.no_fortran:
        build_fortran()
