"""The C-ABI library exists and exports every entry point include/mcx.h
declares (no GPU needed: only symbol lookup)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mcx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mcx_[a-z_]+)\s*\(", src)))


def test_header_declares_the_boundary():
    names = declared_functions()
    for n in ("mcx_create", "mcx_destroy", "mcx_set_tables", "mcx_set_model", "mcx_transport",
              "mcx_get_tallies", "mcx_temperature", "mcx_last_error", "mcx_last_timing"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from mcmax_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "libmcx.so not built (run __graft_entry__.build())"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in declared_functions():
        assert hasattr(lib, n), n
    assert set(declared_functions()) == set(_lib.EXPORTS)


def test_fortran_binding_layouts():
    """The bind(C) types of mcmax_amd/fortran/mcx_binding.f90 have the sizes
    of the C structures (the driver prints them without a GPU call)."""
    import ctypes as C
    import subprocess
    import pytest
    from mcmax_amd import _lib, build
    if not os.path.exists(build.DRIVER):
        pytest.skip("the Fortran caller is not built (no amdflang)")
    out = subprocess.run([build.DRIVER, "--sizes"], capture_output=True, text=True, timeout=60).stdout.split()
    assert out[0] == "sizes"
    sizes = [int(x) for x in out[1:6]]
    assert sizes == [C.sizeof(_lib.Tables), C.sizeof(_lib.Grid), C.sizeof(_lib.Run), C.sizeof(_lib.Tallies),
                     C.sizeof(_lib.Temps)]
