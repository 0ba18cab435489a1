"""The C ABI library (lib3dcompact_hip.so) loads, exports every function
include/cs3d.h declares, and rejects bad arguments on the host side (no GPU
needed for any of these)."""
import ctypes
import os
import re

import pytest

import cs3d

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if not f.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", f)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names |= set(re.findall(r"\b(cs3d_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    assert "cs3d_create" in names and "cs3d_ops_apply" in names and len(names) >= 20


@pytest.mark.parametrize("name", declared_functions())
def test_symbol_exported(hiplib, name):
    assert hasattr(hiplib, name), name


def test_version_and_error_string(hiplib):
    assert b"gfx950" in hiplib.cs3d_version()
    assert hiplib.cs3d_last_error() is not None


def test_config_default_matches_reference(hiplib):
    c = cs3d.config(16, 16, 16)
    assert c.gamma == 1.4 and c.Pr == 0.7 and c.alpha_f == 0.495  # IdealGas.hpp:39-40, 3DCompact.cpp:90
    assert tuple(c.sponge_frac) == (0.25, 0.125, 0.25)  # SpongeBC.hpp:57-59
    assert c.sponge_strength == 6.0 and c.sponge_avgT == 15.0 and c.sponge_epsP == 0.005


def _create(cfg):
    h = ctypes.c_void_p()
    rc = cs3d.lib().cs3d_create(ctypes.byref(cfg), ctypes.byref(h))
    return rc, h


def test_rejects_inlet_and_short_wall_lines_before_touching_the_gpu(hiplib):
    # BC.hpp:11 declares INLET; CSolver_AWS handles no such face
    c = cs3d.config(16, 16, 16, bc_type=(0, 1, 0), bc_kind=(0, 0, 4, 2, 0, 0))
    rc, h = _create(c)
    assert rc == cs3d.ERR_UNSUPPORTED and not h
    assert b"inlet" in hiplib.cs3d_last_error()
    # the wall temperature takes the 6 points next to the wall
    c = cs3d.config(16, 6, 16, bc_type=(0, 1, 0), bc_kind=(0, 0, 2, 2, 0, 0))
    rc, h = _create(c)
    assert rc == cs3d.ERR_ARG and not h
    assert b"wall" in hiplib.cs3d_last_error()


@pytest.mark.parametrize("field,value", [("Nx", 3), ("rk", 2), ("precision", 7),
                                         ("filter_step", 0), ("gamma", 1.0), ("Lx", 0.0)])
def test_rejects_bad_config(hiplib, field, value):
    c = cs3d.config(16, 16, 16)
    setattr(c, field, value)
    rc, h = _create(c)
    assert rc == cs3d.ERR_ARG and not h


def test_null_arguments(hiplib):
    assert hiplib.cs3d_create(None, None) == cs3d.ERR_ARG
    assert hiplib.cs3d_step(None, 1) == cs3d.ERR_ARG
    assert hiplib.cs3d_ops_apply(None, 1, 0, None, None, 0) == cs3d.ERR_ARG


def test_plan_rejects_short_lines(hiplib):
    ints = (ctypes.c_int * 8)()
    # the filter needs 8 points (Filter.cpp:5-96); non-periodic lines need 8
    assert hiplib.cs3d_plan_info(0, 0, 7, 0.1, 0.49, 0, ints, None, 0, None, 0) == cs3d.ERR_ARG
    assert hiplib.cs3d_plan_info(1, 1, 7, 0.1, 0.49, 0, ints, None, 0, None, 0) == cs3d.ERR_ARG
    assert hiplib.cs3d_plan_info(1, 0, 64, 0.1, 0.49, 0, ints, None, 0, None, 0) == cs3d.OK
    assert ints[0] == 64 and ints[4] == 1  # uniform periodic plan
