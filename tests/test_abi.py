"""CPU checks of the C-ABI boundary: the library loads and exports every
entry point include/lcode3d_hip.h declares (no compute: no GPU here)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lcode3d_hip.h")


def _declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(lcode_\w+)\s*\(", text)))


def test_header_declares_the_step_surface():
    names = _declared()
    for n in ("lcode_step", "lcode_deposit", "lcode_move_smart",
              "lcode_move_estimate_wo_fields", "lcode_calculate_Ez",
              "lcode_calculate_Bz", "lcode_calculate_Ex_Ey_Bx_By",
              "lcode_workspace_create", "lcode_workspace_destroy",
              "lcode_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from lcode3d_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail(f"{_lib.LIB_PATH} not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(lib, name), name
    # the Python binding covers exactly the declared surface
    assert sorted(_lib.SYMBOLS) == _declared()


def test_version_and_error_string_without_gpu():
    from lcode3d_amd import _lib
    lib = _lib.load()
    assert b"gfx950" in lib.lcode_version()
    assert isinstance(lib.lcode_last_error(), bytes)


def test_invalid_config_is_rejected_on_the_host():
    """Validation happens before any device call, so it runs here."""
    from lcode3d_amd import _lib
    lib = _lib.load()
    cfg = _lib.LcodeConfig(grid_steps=640, grid_step_size=.025,
                           xi_step_size=.005, subtraction_trick=1.,
                           variant_A=1, reflect_boundary=7.,
                           plasma_coarseness=3, plasma_fineness=2,
                           nc=205, nf=1240, solver=0)
    ws = ctypes.c_void_p()
    rc = lib.lcode_workspace_create(ctypes.byref(cfg), ctypes.byref(ws))
    assert rc == -1
    assert b"odd" in lib.lcode_last_error()
    assert lib.lcode_step(None, ctypes.byref(cfg), None, None, None, None,
                          0., None, None) == -1


def test_product_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "lcode3d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".hpp", ".cpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert not re.search(r"^\s*(from|import)\s+oracle", text,
                                     re.M), f
