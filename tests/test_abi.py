"""CPU-side checks of the C ABI: the library loads and exports every
function include/fdfk.h declares (no GPU calls)."""
import os
import re

from fdfk2d_amd import lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "fdfk.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", txt)
    keep = [n for n in names if n.startswith("fdfk_") or n.endswith("_")]
    return sorted(set(keep))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("fdfk_create", "fdfk_run", "fdfk_seismograms", "updatedispl_",
              "comput_background_wavefield_"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = lib.lib()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_library_is_built_for_gfx950():
    data = open(lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data


def test_fortran_binding_names_exported_and_declared():
    """Every bind(C, name=...) of the Fortran module include/fdfk.f90 is a
    function the header declares and the library exports."""
    txt = open(os.path.join(ROOT, "include", "fdfk.f90")).read()
    names = re.findall(r'bind\(C,\s*name="([a-z_0-9]+)"\)', txt)
    assert len(names) >= 8
    L = lib.lib()
    declared = declared_functions()
    assert not [n for n in names if n not in declared or not hasattr(L, n)]


def test_library_resolves_every_symbol_at_load():
    """Loading with RTLD_NOW binds every undefined symbol at once: a
    declared-but-undefined internal function fails here, not on first use."""
    import ctypes
    ctypes.CDLL(lib.LIB_PATH, mode=os.RTLD_NOW | os.RTLD_LOCAL)
