"""The C ABI library loads and exports every symbol include/vacgpu.h declares.

Runs without a GPU: only dlopen and symbol lookup, no device call.
"""
import os
import re

import pytest

from vac_amd import vacgpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vacgpu.h")).read()
    return sorted(set(re.findall(r"\b(vacgpu_[a-z_]+)\s*\(", text)))


def test_header_declares_the_binding_signatures():
    assert declared_symbols() == sorted(vacgpu.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(vacgpu.LIB_PATH):
        pytest.fail("libvacgpu.so not built (python -m vac_amd.build)")
    lib = vacgpu.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_supported_rejects_unsupported_selectors():
    lib = vacgpu.load_library()
    import ctypes as C
    cfg = vacgpu.Config()
    cfg.ndim = 2
    cfg.physics = vacgpu.HD
    cfg.nx[0] = cfg.nx[1] = 16
    cfg.eqpar[0] = 1.4
    m = vacgpu.MethodSpec(typefull="tvd").to_c()
    assert lib.vacgpu_supported(C.byref(cfg), C.byref(m)) == 0  # HD Roe solver
    # a typeB the path does not implement: rejected with an error, not a fallback
    cfg.typeB[0][0] = 9  # ('special' and the others: no code)
    assert lib.vacgpu_supported(C.byref(cfg), C.byref(m)) == -1
    cfg.typeB[0][0] = vacgpu.CONT
    m2 = vacgpu.MethodSpec(typefull="tvdlf").to_c()
    assert lib.vacgpu_supported(C.byref(cfg), C.byref(m2)) == 0
    cfg.physics = vacgpu.MHD
    cfg.eqpar[1] = 0.1  # resistive MHD is not on the path
    assert lib.vacgpu_supported(C.byref(cfg), C.byref(m2)) == -1
