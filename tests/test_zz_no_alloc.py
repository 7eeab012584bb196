"""vacgpu_getdt_courant_noalloc: getdt_courant on a caller's device array with
no grid handle and no allocation (its partials and scalars live in a fixed
scratch array of the library), against vacgpu_getdt_courant on a grid
holding the same state.  dt and the pressure floor must be the same doubles
(the kernels and the order-free max / min folds are the same), call after
call, each call smaller than the one before it, so that what a call leaves
in the shared scratch cannot leak into the next.
"""
import ctypes as C

import numpy as np
import pytest

from vac_amd import vacgpu
from vac_amd.grid import widths_1d
from vac_amd.initial import random_fourier

pytestmark = pytest.mark.gpu


class _DeviceArray:
    """A device copy of a host array, through the HIP runtime the library
    links (the test allocates; the entry point under test does not)."""

    _hip = None

    def __init__(self, a):
        if _DeviceArray._hip is None:
            _DeviceArray._hip = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
        hip = _DeviceArray._hip
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert hip.hipMalloc(C.byref(self.ptr), C.c_size_t(a.nbytes)) == 0
        assert hip.hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes),
                             1) == 0  # hipMemcpyHostToDevice

    def free(self):
        if self.ptr:
            _DeviceArray._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _device(a):
    return _DeviceArray(a)


def _expected(w_mesh, coords, physics, gamma, method, dtcourant=None):
    nx = [len(c) for c in coords]
    g = vacgpu.Grid(nx, physics, gamma, [widths_1d(c) for c in coords])
    g.put_mesh(w_mesh)
    w_full = g.get_state()
    dt, _, sp = g.getdt_courant(method, dtcourant=dtcourant, smallp=-1.0)
    g.close()
    return w_full, dt, sp


def _noalloc(w_full, coords, physics, gamma, method, pad=3, dtcourant=None, smallp=-1.0):
    """Put w_full into a device array with leading dimension nx+4+pad along x
    and call the entry point through the C ABI."""
    lib = vacgpu.load_library()
    nd = len(coords)
    nx = [len(c) for c in coords]
    ng = [n + 4 for n in nx]
    ld = [ng[0] + pad] + ng[1:] + [1] * (3 - nd)
    host = np.zeros((w_full.shape[0],) + tuple(reversed(ld[:nd])))
    host[(slice(None),) + tuple(slice(0, n) for n in reversed(ng))] = w_full
    w_dev = _device(host)
    dx_dev = [_device(widths_1d(c)) for c in coords]
    cfg = vacgpu.Config()
    cfg.ndim = nd
    cfg.physics = vacgpu.MHD if physics == "mhd" else vacgpu.HD
    for d in range(nd):
        cfg.nx[d] = nx[d]
    cfg.eqpar[0] = gamma
    m = method.to_c(cfg.physics)
    ldc = (C.c_int * 3)(*ld)
    dtc = None if dtcourant is None else np.ascontiguousarray(dtcourant, dtype=np.float64)
    sp = C.c_double(smallp)
    dt = C.c_double()
    err = C.create_string_buffer(256)
    dxp = [t.ptr for t in dx_dev] + [None] * (3 - nd)
    rc = lib.vacgpu_getdt_courant_noalloc(
        C.byref(cfg), C.byref(m), w_dev.ptr, ldc, dxp[0], dxp[1], dxp[2],
        dtc.ctypes.data_as(C.POINTER(C.c_double)) if dtc is not None else None,
        C.byref(sp), C.byref(dt), None, err, 256)
    for t in [w_dev] + dx_dev:
        t.free()
    return rc, dt.value, sp.value, dtc, err.value.decode()


def test_noalloc_getdt_matches_the_grid_entry_point():
    cases = [
        # (shape, seed, physics, gamma, courantpar, lagged dtcourant), each smaller than the last
        ((96, 80), 901, "mhd", 5 / 3, 0.8, None),
        ((40, 36), 902, "mhd", 1.4, 0.6, None),
        ((20, 16, 12), 903, "mhd", 5 / 3, 0.6, None),
        ((16, 12, 10), 904, "hd", 1.4, 0.8, None),
        ((12, 10), 905, "mhd", 5 / 3, 0.8, [1e99, 3.5e-3]),
    ]
    for shape, seed, physics, gamma, cp, lag in cases:
        print(f"case {shape} {physics} courantpar={cp} lagged={lag}", flush=True)
        w0, coords = random_fourier(shape, seed, gamma, physics=physics)
        method = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=cp)
        w_full, dt_ref, sp_ref = _expected(w0, coords, physics, gamma, method, dtcourant=lag)
        rc, dt, sp, dtc, err = _noalloc(w_full, coords, physics, gamma, method, dtcourant=lag)
        assert rc == 0, err
        assert dt == dt_ref, (dt, dt_ref)
        assert sp == sp_ref, (sp, sp_ref)
        if lag is not None:
            assert list(dtc) == [1e99] * len(lag)  # reset as VAC's getdt_courant does
        # a given pressure floor is kept, and dt does not depend on it here
        rc, dt2, sp2, _, err = _noalloc(w_full, coords, physics, gamma, method, pad=0,
                                        dtcourant=lag, smallp=sp_ref)
        assert rc == 0, err
        assert (dt2, sp2) == (dt_ref, sp_ref)


def test_noalloc_rejects_a_grid_beyond_its_scratch():
    # the size check comes before any HIP call, so no pointer is touched
    lib = vacgpu.load_library()
    cfg = vacgpu.Config()
    cfg.ndim = 2
    cfg.physics = vacgpu.MHD
    cfg.nx[0], cfg.nx[1] = 8192, 4096
    cfg.eqpar[0] = 5 / 3
    m = vacgpu.MethodSpec(typefull="tvdlf", courantpar=0.8).to_c(vacgpu.MHD)
    ld = (C.c_int * 3)(8196, 4100, 1)
    sp, dt = C.c_double(-1.0), C.c_double()
    err = C.create_string_buffer(256)
    unused = C.c_void_p(16)
    rc = lib.vacgpu_getdt_courant_noalloc(C.byref(cfg), C.byref(m), unused, ld, unused, unused,
                                          None, None, C.byref(sp), C.byref(dt), None, err, 256)
    assert rc == -1
    assert "too large for the fixed scratch" in err.value.decode()
