"""CPU checks of the C ABI library: it loads and exports every symbol that
include/libheom_amd.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import libheom_amd
from libheom_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "libheom_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(heom_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_signatures_load():
    L = _capi.load()
    major, minor = ctypes.c_int(), ctypes.c_int()
    assert L.heom_version(ctypes.byref(major), ctypes.byref(minor)) == 0
    opt = _capi.heom_options()
    assert L.heom_default_options(ctypes.byref(opt)) == 0
    assert (opt.parity_q1, opt.sqrt_norm, opt.order, opt.device) == (1, 0, 0, 0)


def test_null_arguments_are_rejected_without_a_device():
    L = _capi.load()
    assert L.heom_create(0, 0, 2, 2, None, None) == _capi.HEOM_ERR_ARG
    assert b"out is NULL" in L.heom_last_error()
    h = ctypes.c_void_p()
    assert L.heom_create(7, 0, 2, 2, None, ctypes.byref(h)) == _capi.HEOM_ERR_ARG
    assert L.heom_set_param(None) == _capi.HEOM_ERR_ARG
    assert L.heom_solver_create(None, 0, None) == _capi.HEOM_ERR_ARG


def test_package_mirrors_reference_class_names():
    for name in ("heom_hilb", "heom_liou", "heom_ado", "rk4", "rkdp", "qme_solver", "hip"):
        assert hasattr(libheom_amd, name)
