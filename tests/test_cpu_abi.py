"""CPU-side checks of the boundary: the C-ABI library loads and exports every
symbol include/adc2g.h declares (no compute call without a GPU), the Python
binding covers the same set, the Fortran iso_c_binding module declares them,
and the Fortran shim compiles with the image's flang."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "adc2g.h")
LIB = os.path.join(ROOT, "adc_ottawa_amd", "libadc2g.so")


def declared():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(adc2g_[a-z_0-9]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("adc2g_create", "adc2g_destroy", "adc2g_sigma", "adc2g_davidson", "adc2g_sil_autocorr",
                 "adc2g_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libadc2g.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    from adc_ottawa_amd import lib
    assert sorted(lib.SYMBOLS) == declared()


def test_fortran_module_binds_core_entry_points():
    src = open(os.path.join(ROOT, "adc_ottawa_amd", "fortran", "adc2g_mod.f90")).read()
    for n in ("adc2g_create", "adc2g_destroy", "adc2g_dims", "adc2g_diagonal", "adc2g_sigma",
              "adc2g_matxvec_treal", "adc2g_davidson", "adc2g_sil_autocorr", "adc2g_last_error"):
        assert f"name='{n}'" in src, n


@pytest.mark.skipif(shutil.which("/opt/rocm/bin/amdflang") is None, reason="no flang")
def test_fortran_module_compiles(tmp_path):
    mod = os.path.join(ROOT, "adc_ottawa_amd", "fortran", "adc2g_mod.f90")
    r = subprocess.run(["/opt/rocm/bin/amdflang", "-c", "-J", str(tmp_path), mod, "-o",
                        str(tmp_path / "m.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_error_codes_match_reference_sil_codes():
    src = open(HEADER).read()
    # source/propagation/sillib.f90:84-89 error codes 1..4
    for name, code in [("ADC2G_ESIL_ORDER", 1), ("ADC2G_ESIL_UNDERFLOW", 2), ("ADC2G_ESIL_DIAG", 3),
                       ("ADC2G_ESIL_RESTART", 4)]:
        assert re.search(rf"#define {name} {code}\b", src)


DRIVER = os.path.join(ROOT, "adc_ottawa_amd", "adc2g_driver")


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="Fortran driver not built")
def test_fortran_driver_rejects_ncore_beyond_core_array(tmp_path):
    # the namelist's core(16) holds at most 16 orbitals: a larger ncore stops
    # with a message before any file is read or any library call is made
    inp = "&adc2g nocc=20 nvirt=40 naux=10 ncore=17 /\n"
    r = subprocess.run([DRIVER], input=inp, capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode != 0
    assert "ncore must be between 0 and 16" in r.stdout
