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


def test_fvm_entry_points_reject_bad_operators_without_gpu():
    """The fully implicit FVM entry points check the operator before any device work: an odd
    ordinate count, a stencil wider than 8, a missing work buffer or aliased input/output
    return BZ_EINVAL (-1); the struct layout matches the header's."""
    from boltzmann_amd import _lib
    L = _lib.lib()
    op = _lib.Fvm1d3v()
    op.Nr, op.Nvt, op.Nx, op.nvr, op.nvt, op.nx = 40, 8, 10, 4, 4, 2
    p = ctypes.cast(ctypes.pointer(op), ctypes.c_void_p)
    buf = ctypes.c_void_p(0x1000)                       # never dereferenced: rejected first
    assert L.bz_fvm_rhs_mvec(p, buf, 1e-3, buf, ctypes.c_void_p(0x2000), None) == -1   # no work buffer
    op.work = 0x3000
    assert L.bz_fvm_rhs_mvec(p, buf, 1e-3, buf, buf, None) == -1                       # x aliases y
    op.Nvt = 9
    assert L.bz_fvm_precond_xvtvrc(p, buf, 1e-3, buf, ctypes.c_void_p(0x2000), None) == -1
    op.Nvt, op.nvr = 8, 9
    assert L.bz_fvm_precond_xvtvrc(p, buf, 1e-3, buf, ctypes.c_void_p(0x2000), None) == -1
    assert L.bz_fvm_rhs_mvec(None, buf, 1e-3, buf, ctypes.c_void_p(0x2000), None) == -1
    # 6 ints, 15 pointers, 5 doubles, 11 pointers
    assert ctypes.sizeof(_lib.Fvm1d3v) == 6 * 4 + 15 * 8 + 5 * 8 + 11 * 8
