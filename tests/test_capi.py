"""The C-ABI libraries load without a GPU and export every symbol include/sogl.h declares."""
import ctypes
import os

import sogl

HOST_ONLY = {"sogl_mtx_read", "sogl_mtx_read_rhs", "sogl_mtx_info", "sogl_mtx_arrays", "sogl_mtx_free",
             "sogl_last_error"}


def test_header_symbols_exported():
    syms = sogl.exported_symbols()
    assert "sogl_factor_run" in syms and "sogl_solve" in syms and "sogl_plan_create" in syms
    lib = ctypes.CDLL(os.path.join(sogl.BUILD, "libsogl.so"))
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_host_library_has_plan_symbols():
    lib = ctypes.CDLL(os.path.join(sogl.BUILD, "libsogl_host.so"))
    for s in sogl.exported_symbols():
        if s.startswith(("sogl_plan", "sogl_mtx")) or s in HOST_ONLY:
            assert hasattr(lib, s), s


def test_plan_errors_are_codes_not_crashes():
    lib = sogl.load("libsogl_host.so")
    h = ctypes.c_void_p()
    assert lib.sogl_plan_create(0, 0, 0, None, None, None, None, 0, ctypes.byref(h)) == 1
    assert lib.sogl_last_error()
