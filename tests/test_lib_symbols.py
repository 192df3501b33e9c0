"""libbz.so loads on the CPU-only host and exports every entry point declared in include/bz.h."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "bz.h")).read()
    return sorted(set(re.findall(r"^int\s+(bz_\w+)\s*\(", txt, flags=re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "bz_steady_solve_0d" in syms and "bz_collision_radial" in syms


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "boltzmann_amd", "libbz.so"))
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_signatures_cover_header():
    from boltzmann_amd import _lib
    _lib.lib()
    missing = [s for s in declared_symbols() if s not in _lib._SIGS]
    assert not missing, missing


def test_window_query_without_gpu():
    lib = ctypes.CDLL(os.path.join(ROOT, "boltzmann_amd", "libbz.so"))
    assert lib.bz_solver0d_window(10) == 11
    assert lib.bz_solver0d_window(7) == 8
    assert lib.bz_solver0d_window(100) == -1
