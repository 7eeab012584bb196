"""Python host side of the MI355X advect path, over the C ABI in include/vacgpu.h.

This mirrors the pieces of VAC's host loop that call the path:
``Grid.getdt_courant`` (src/vac.t:664-722), ``Grid.advect`` (src/vac.t:363-486)
and ``Grid.run`` (the time loop of src/vac.t:55-106).  The state stays on the
device between calls; ``put_state`` / ``get_state`` move the full ixG array.

The shared library is built in-tree by ``python -m vac_amd.build``; importing
this module on a machine without it raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvacgpu.so")

HD, MHD = 0, 1
TVDLF, TVD, TVDMU = 0, 1, 2
LIMITERS = {"minmod": 0, "woodward": 1, "superbee": 2}
ENTROPY = {"nul": 0, "harten": 1, "powell": 2, "ratio": 3, "yee": 4}
CONT, PERIODIC, SYMM, ASYMM, FIXED, CONT1, FIXED1, GRAD1, SYMM0 = 0, 1, 2, 3, 4, 5, 6, 7, 8
BC_CODES = {"cont": CONT, "periodic": PERIODIC, "symm": SYMM, "asymm": ASYMM, "fixed": FIXED,
            "cont1": CONT1, "fixed1": FIXED1, "grad1": GRAD1, "symm0": SYMM0}
FP64, FP32BULK = 0, 1
KERNELS = {"tvd_sweep": 0, "tvdlf_pred": 1, "tvdlf_corr": 2, "ryu": 3, "bound": 4, "reduce": 5,
           "tvdlf_fused": 6, "projectb": 7}


class Config(C.Structure):
    _fields_ = [("ndim", C.c_int), ("physics", C.c_int), ("nx", C.c_int * 3),
                ("eqpar", C.c_double * 2), ("typeB", (C.c_int * 8) * 6),
                ("precision", C.c_int)]


class Method(C.Structure):
    _fields_ = [("typefull", C.c_int), ("typelimiter", C.c_int * 8), ("divbfix", C.c_int),
                ("divbwave", C.c_int), ("sourceunsplit", C.c_int), ("dimsplit_xy", C.c_int),
                ("courantpar", C.c_double), ("smallpcoeff", C.c_double),
                ("typelimited", C.c_int), ("typeentropy", C.c_int * 8),
                ("entropycoef", C.c_double * 8), ("useprimitive", C.c_int),
                ("unsplit", C.c_int), ("dtcantgrow", C.c_int), ("slowsteps", C.c_int)]


_lib = None
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)

# every entry point declared in include/vacgpu.h with its ctypes signature
SIGNATURES = {
    "vacgpu_create": ([C.POINTER(Config), _DP, _DP, _DP, C.POINTER(C.c_void_p)], C.c_int),
    "vacgpu_set_stream": ([C.c_void_p, C.c_void_p], C.c_int),
    "vacgpu_put_state": ([C.c_void_p, _DP, _IP], C.c_int),
    "vacgpu_put_state_ghosts": ([C.c_void_p, _DP, _IP], C.c_int),
    "vacgpu_get_state": ([C.c_void_p, _DP, _IP], C.c_int),
    "vacgpu_getdt_courant": ([C.c_void_p, C.POINTER(Method), _DP, _DP, _DP], C.c_int),
    "vacgpu_advect": ([C.c_void_p, C.POINTER(Method), C.c_int, C.c_int, C.c_double, C.c_int,
                       C.c_int, _DP, _IP, _DP], C.c_int),
    "vacgpu_run": ([C.c_void_p, C.POINTER(Method), C.c_int, C.c_int, C.c_double, C.c_int,
                    C.c_double, _DP], C.c_int),
    "vacgpu_set_time": ([C.c_void_p, C.c_int, C.c_double], C.c_int),
    "vacgpu_get_time": ([C.c_void_p, _IP, _DP], C.c_int),
    "vacgpu_get_scalars": ([C.c_void_p, _DP, _IP, _DP, _DP], C.c_int),
    "vacgpu_profile": ([C.c_void_p, C.c_int], C.c_int),
    "vacgpu_kernel_time": ([C.c_void_p, C.c_int, _DP, C.POINTER(C.c_long)], C.c_int),
    "vacgpu_set_kernels": ([C.c_void_p, C.c_int, C.c_int], C.c_int),
    "vacgpu_check_quotients": ([_DP, _DP, C.c_longlong, _DP, _DP], C.c_int),
    "vacgpu_check_roots": ([_DP, C.c_longlong, _DP, _DP], C.c_int),
    "vacgpu_snapshot_start": ([C.c_void_p, C.c_int], C.c_int),
    "vacgpu_snapshot_fetch": ([C.c_void_p, C.c_int, _DP, _IP, _IP, _DP], C.c_int),
    "vacgpu_snapshot_fetch_mesh": ([C.c_void_p, C.c_int, _DP, _IP, _DP], C.c_int),
    "vacgpu_set_process": ([C.c_void_p, C.c_int, _DP, C.c_int], C.c_int),
    "vacgpu_process": ([C.c_void_p, C.c_int, _DP, _IP, _IP, _DP], C.c_int),
    "vacgpu_getdt_courant_noalloc": ([C.POINTER(Config), C.POINTER(Method), C.c_void_p, _IP,
                                      C.c_void_p, C.c_void_p, C.c_void_p, _DP, _DP, _DP,
                                      C.c_void_p, C.c_char_p, C.c_int], C.c_int),
    "vacgpu_supported": ([C.POINTER(Config), C.POINTER(Method)], C.c_int),
    "vacgpu_last_error": ([C.c_void_p], C.c_char_p),
    "vacgpu_destroy": ([C.c_void_p], None),
    "vacgpu_batch_footprint": ([C.POINTER(Config)], C.c_longlong),
    "vacgpu_batch_create": ([C.c_longlong, C.c_int, C.POINTER(C.c_void_p)], C.c_int),
    "vacgpu_batch_add": ([C.c_void_p, C.POINTER(Config), C.POINTER(Method), _DP, _DP, _DP, _IP,
                          C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, _IP], C.c_int),
    "vacgpu_batch_remove": ([C.c_void_p, C.c_int], C.c_int),
    "vacgpu_batch_step": ([C.c_void_p, C.c_int], C.c_int),
    "vacgpu_batch_sync": ([C.c_void_p], C.c_int),
    "vacgpu_batch_status": ([C.c_void_p, C.c_int, _IP, _DP, _DP, _IP, _IP], C.c_int),
    "vacgpu_batch_get_state": ([C.c_void_p, C.c_int, _DP, _IP], C.c_int),
    "vacgpu_batch_snapshot": ([C.c_void_p], C.c_int),
    "vacgpu_batch_snapshot_wait": ([C.c_void_p, C.c_int], C.c_int),
    "vacgpu_batch_fetch_start": ([C.c_void_p, C.c_int, C.c_int, _IP], C.c_int),
    "vacgpu_batch_finished": ([C.c_void_p, _IP, C.c_int], C.c_int),
    "vacgpu_batch_add_mesh": ([C.c_void_p, C.POINTER(Config), C.POINTER(Method), _DP, _DP, _DP,
                               C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, _IP], C.c_int),
    "vacgpu_batch_fetch_ready": ([C.c_void_p, C.c_int, _DP, _IP, C.c_int], C.c_int),
    "vacgpu_batch_graph": ([C.c_void_p, C.c_int], C.c_int),
    "vacgpu_batch_tune": ([C.c_void_p, _IP], C.c_int),
    "vacgpu_batch_set_tiles": ([C.c_void_p, _IP], C.c_int),
    "vacgpu_batch_set_kernels": ([C.c_void_p, C.c_int, _IP], C.c_int),
    "vacgpu_batch_get_tiles": ([C.c_void_p, _IP, _IP], C.c_int),
    "vacgpu_batch_bytes_used": ([C.c_void_p], C.c_longlong),
    "vacgpu_batch_last_error": ([C.c_void_p], C.c_char_p),
    "vacgpu_batch_destroy": ([C.c_void_p], None),
}


def load_library(path: str = LIB_PATH):
    """Load libvacgpu.so and declare its signatures (no device call is made)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m vac_amd.build` "
                           "(the advect path has no CPU fallback)")
    L = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a):
    return a.ctypes.data_as(_DP) if a is not None else None


class VacGpuError(RuntimeError):
    pass


@dataclass
class MethodSpec:
    typefull: str = "tvdlf"
    limiter: object = "minmod"          # one name, or a list per variable / wave
    divbfix: bool = True
    divbwave: bool = True
    sourceunsplit: bool = True
    dimsplit_xy: bool = False
    courantpar: float = 0.8
    smallpcoeff: float = 1e-6
    typelimited: str = "original"       # or "previous", "predictor"
    typeentropy: object = "nul"         # one name, or a list per wave
    entropycoef: object = None          # per wave; None: VAC's defaults
    useprimitive: bool = False          # limit primitive variables (FP64)
    unsplit: bool = False               # dimsplit=F (TVDLF, TVD-MUSCL)
    dtcantgrow: bool = False            # the run's dt may not grow (src/vac.t:68-69)
    slowsteps: int = 0                  # slow start over the first steps (src/vac.t:75-78)

    def to_c(self, physics: str = "mhd") -> Method:
        m = Method()
        m.typefull = {"tvdlf": TVDLF, "tvd": TVD, "tvdmu": TVDMU}[self.typefull]
        lims = [self.limiter] * 8 if isinstance(self.limiter, str) else list(self.limiter) + ["minmod"] * 8
        for i in range(8):
            m.typelimiter[i] = LIMITERS[lims[i]]
        m.divbfix = int(self.divbfix)
        m.divbwave = int(self.divbwave)
        m.sourceunsplit = int(self.sourceunsplit)
        m.dimsplit_xy = int(self.dimsplit_xy)
        m.courantpar = self.courantpar
        m.smallpcoeff = self.smallpcoeff
        m.typelimited = {"original": 0, "previous": 1, "predictor": 2}[self.typelimited]
        ents = [self.typeentropy] * 8 if isinstance(self.typeentropy, str) else \
            list(self.typeentropy) + ["nul"] * 8
        for i in range(8):
            m.typeentropy[i] = ENTROPY[ents[i]]
        coef = self.entropycoef
        if coef is None:
            # physini's defaults (0.2 fast/slow/sound, 0.4 Alfven, -1 else;
            # single-precision literals stored as doubles)
            f2, f4 = float(np.float32(0.2)), float(np.float32(0.4))
            coef = [f2, f2, f2, f2, -1.0, -1.0, f4, f4]
            if physics == "hd":
                coef = [f2, f2] + [-1.0] * 6
        coef = list(coef) + [-1.0] * 8
        for i in range(8):
            m.entropycoef[i] = float(coef[i])
        m.useprimitive = int(self.useprimitive)
        m.unsplit = int(self.unsplit)
        m.dtcantgrow = int(self.dtcantgrow)
        m.slowsteps = int(self.slowsteps)
        return m


def nvars(physics: str, ndim: int) -> int:
    return 2 + 2 * ndim if physics == "mhd" else 2 + ndim


def make_config(nx, physics: str, gamma: float, bc="periodic", precision: str = "fp64") -> Config:
    ndim = len(nx)
    cfg = Config()
    cfg.ndim = ndim
    cfg.physics = MHD if physics == "mhd" else HD
    for d in range(ndim):
        cfg.nx[d] = int(nx[d])
    cfg.eqpar[0] = gamma
    cfg.eqpar[1] = 0.0
    # bc: one type for all faces, or one entry per face (x-low, x-high,
    # y-low, ...); an entry is a type or a list of types per variable
    bcs = [bc] * (2 * ndim) if isinstance(bc, str) else list(bc)
    for f in range(2 * ndim):
        spec = bcs[f]
        for iw in range(8):
            name = spec if isinstance(spec, str) else spec[min(iw, len(spec) - 1)]
            if name not in BC_CODES:
                raise ValueError(f"typeB {name!r} is not on the GPU path")
            cfg.typeB[f][iw] = BC_CODES[name]
    cfg.precision = FP32BULK if precision == "fp32bulk" else FP64
    return cfg


class Grid:
    """One grid resident on the GPU (a vacgpu_grid handle)."""

    def __init__(self, nx, physics: str, gamma: float, dxs, bc="periodic",
                 precision: str = "fp64"):
        self.lib = load_library()
        self.nx = [int(n) for n in nx]
        self.ndim = len(self.nx)
        self.physics = physics
        self.nw = nvars(physics, self.ndim)
        cfg = make_config(self.nx, physics, gamma, bc=bc, precision=precision)
        self.cfg = cfg
        self._dx = [np.ascontiguousarray(d, dtype=np.float64) for d in dxs]
        for d in range(self.ndim):
            if self._dx[d].shape != (self.nx[d] + 4,):
                raise ValueError("dx arrays must have nx+4 entries (ixG)")
        h = C.c_void_p()
        dx = self._dx + [None] * (3 - self.ndim)
        rc = self.lib.vacgpu_create(C.byref(cfg), _ptr(dx[0]), _ptr(dx[1]), _ptr(dx[2]), C.byref(h))
        if rc != 0:
            raise VacGpuError(f"vacgpu_create failed ({rc})")
        self.h = h

    # ------------------------------------------------------------------
    def _check(self, rc, what):
        if rc < 0:
            msg = self.lib.vacgpu_last_error(self.h)
            raise VacGpuError(f"{what}: {msg.decode() if msg else rc}")
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.lib.vacgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape_full(self):
        return (self.nw,) + tuple(n + 4 for n in reversed(self.nx))

    def put_state(self, w_full: np.ndarray):
        w = np.ascontiguousarray(w_full, dtype=np.float64)
        if w.shape != self.shape_full:
            raise ValueError(f"state shape {w.shape} != {self.shape_full}")
        self._check(self.lib.vacgpu_put_state(self.h, _ptr(w), None), "put_state")

    def put_state_ghosts(self, w_full: np.ndarray):
        """Put a full ixG state keeping its ghost cells as they are."""
        w = np.ascontiguousarray(w_full, dtype=np.float64)
        if w.shape != self.shape_full:
            raise ValueError(f"state shape {w.shape} != {self.shape_full}")
        self._check(self.lib.vacgpu_put_state_ghosts(self.h, _ptr(w), None), "put_state_ghosts")

    def put_mesh(self, w_mesh: np.ndarray):
        """Put a mesh-only state; ghost cells are filled by the first call."""
        pad = [(0, 0)] + [(2, 2)] * self.ndim
        self.put_state(np.pad(np.asarray(w_mesh, dtype=np.float64), pad))

    def get_state(self) -> np.ndarray:
        w = np.empty(self.shape_full)
        self._check(self.lib.vacgpu_get_state(self.h, _ptr(w), None), "get_state")
        return w

    def get_mesh(self) -> np.ndarray:
        w = self.get_state()
        return np.ascontiguousarray(w[(slice(None),) + (slice(2, -2),) * self.ndim])

    def snapshot_start(self, slot: int = 0):
        """Start an asynchronous snapshot of the current state into `slot`
        (0 or 1); returns at once, later steps may follow."""
        self._check(self.lib.vacgpu_snapshot_start(self.h, slot), "snapshot_start")

    def snapshot_fetch(self, slot: int = 0, mesh: bool = True):
        """Wait for the snapshot in `slot`; returns (w, it, t) with w the mesh
        (or, mesh=False, the whole ixG array) as get_mesh / get_state give it."""
        it = C.c_int()
        t = C.c_double()
        if mesh:
            w = np.empty((self.nw,) + tuple(reversed(self.nx)))
            self._check(self.lib.vacgpu_snapshot_fetch_mesh(self.h, slot, _ptr(w), C.byref(it),
                                                            C.byref(t)), "snapshot_fetch")
        else:
            w = np.empty(self.shape_full)
            self._check(self.lib.vacgpu_snapshot_fetch(self.h, slot, _ptr(w), None, C.byref(it),
                                                       C.byref(t)), "snapshot_fetch")
        return w, it.value, t.value

    def set_process(self, nproc1: int, procpar=(-1.0, -1.0, -1.0), itmin: int | None = None):
        """VAC's nproc(1) and procpar: project B before the steps (or sweeps)
        VAC's advect would; itmin defaults to the current it."""
        if itmin is None:
            itmin = self.get_time()[0]
        pp = (C.c_double * 3)(*[float(v) for v in procpar])
        self._check(self.lib.vacgpu_set_process(self.h, int(nproc1), pp, int(itmin)),
                    "set_process")

    def process(self, count: int = 1, procpar=None):
        """One process(count) call (the projection of B); returns the
        solve's (info, iterations, residual)."""
        pp = None if procpar is None else (C.c_double * 3)(*[float(v) for v in procpar])
        info, it = C.c_int(), C.c_int()
        res = C.c_double()
        self._check(self.lib.vacgpu_process(self.h, int(count), pp, C.byref(info), C.byref(it),
                                            C.byref(res)), "process")
        return info.value, it.value, res.value

    def set_time(self, it: int, t: float):
        self._check(self.lib.vacgpu_set_time(self.h, it, t), "set_time")

    def get_time(self):
        it = C.c_int()
        t = C.c_double()
        self._check(self.lib.vacgpu_get_time(self.h, C.byref(it), C.byref(t)), "get_time")
        return it.value, t.value

    def scalars(self):
        dtc = np.zeros(3)
        ne = np.zeros(4, dtype=np.int32)
        sp = C.c_double()
        dt = C.c_double()
        self._check(self.lib.vacgpu_get_scalars(self.h, _ptr(dtc), ne.ctypes.data_as(_IP),
                                                C.byref(sp), C.byref(dt)), "get_scalars")
        return {"dtcourant": dtc[: self.ndim], "nerror": ne, "smallp": sp.value, "dt": dt.value}

    def getdt_courant(self, method: MethodSpec, dtcourant=None, smallp=None):
        m = method.to_c(self.physics)
        dt = C.c_double()
        dtc = None if dtcourant is None else np.ascontiguousarray(dtcourant, dtype=np.float64)
        sp = None if smallp is None else C.c_double(smallp)
        self._check(self.lib.vacgpu_getdt_courant(self.h, C.byref(m), _ptr(dtc),
                                                  C.byref(sp) if sp is not None else None,
                                                  C.byref(dt)), "getdt_courant")
        return dt.value, dtc, (sp.value if sp is not None else None)

    def advect(self, method: MethodSpec, idim: int, dt: float, firstsweep: bool,
               lastsweep: bool, dtcourant=None, nerror=None, smallp=None, idimmax=None):
        """One sweep along idim (1-based), or, with idimmax, the unsplit
        sweep over idim..idimmax (= 1..ndim)."""
        m = method.to_c(self.physics)
        ne = None if nerror is None else nerror.ctypes.data_as(_IP)
        sp = None if smallp is None else C.c_double(smallp)
        i1 = idim if idimmax is None else idimmax
        self._check(self.lib.vacgpu_advect(self.h, C.byref(m), idim, i1, dt, int(firstsweep),
                                           int(lastsweep), _ptr(dtcourant), ne,
                                           C.byref(sp) if sp is not None else None), "advect")
        return sp.value if sp is not None else None

    def run(self, method: MethodSpec, nsteps: int, itmax: int = 1 << 30, tmax: float = 1e99,
            tmaxexact: bool = False, dtpar: float = -1.0, record_t: bool = False):
        m = method.to_c(self.physics)
        hist = np.zeros(nsteps + 1) if record_t else None
        n = self._check(self.lib.vacgpu_run(self.h, C.byref(m), nsteps, itmax, tmax,
                                            int(tmaxexact), dtpar, _ptr(hist)), "run")
        return n, (hist[: n + 1] if record_t else None)

    def profile(self, enable: bool):
        self._check(self.lib.vacgpu_profile(self.h, int(enable)), "profile")

    def set_kernels(self, fused: bool = True, seglen: int = 0):
        """2D TVDLF sweeps as one fused launch (default) or predictor +
        corrector launches; seglen = rows per marching segment (0 = auto)."""
        self._check(self.lib.vacgpu_set_kernels(self.h, int(fused), int(seglen)), "set_kernels")

    def kernel_time(self, name: str):
        ms = C.c_double()
        cnt = C.c_long()
        self._check(self.lib.vacgpu_kernel_time(self.h, KERNELS[name], C.byref(ms), C.byref(cnt)),
                    "kernel_time")
        return ms.value, cnt.value


def host_loop(grid: Grid, method: MethodSpec, nsteps: int, it0: int = 0, t0: float = 0.0):
    """VAC's time loop (src/vac.t:55-106) driven from the host through
    getdt_courant and advect, with dtcourant / nerror kept on the host as
    VAC's COMMON variables are.  Returns (t history, nerror)."""
    ndim = grid.ndim
    dtcourant = np.full(ndim, 1e99)
    nerror = np.zeros(4, dtype=np.int32)
    smallp = -1.0
    it, t = it0, t0
    hist = [t]
    for _ in range(nsteps):
        dt, dtcourant, smallp = grid.getdt_courant(method, dtcourant, smallp)
        if method.unsplit:  # all directions in one advect (src/vac.t:325-329)
            grid.advect(method, 1, dt, True, True, dtcourant, nerror, smallp, idimmax=ndim)
        order = [] if method.unsplit else \
            list(range(ndim)) if (it % 2 == 0 or method.dimsplit_xy) else list(range(ndim - 1, -1, -1))
        for k, d in enumerate(order):
            grid.advect(method, d + 1, dt, k == 0, k == ndim - 1, dtcourant, nerror, smallp)
        it += 1
        t = t + dt
        hist.append(t)
    return np.array(hist), nerror


class Batch:
    """Many independent 2D TVDLF grids advanced together (vacgpu_batch).

    Each instance follows VAC's time loop with its own dt, t, it and stop
    condition; a finished instance stops on the device.  ``add`` raises
    ``BudgetFull`` when the instance does not fit the HBM budget.
    """

    class BudgetFull(Exception):
        pass

    def __init__(self, hbm_budget_bytes: int, capacity: int = 1024):
        self.lib = load_library()
        h = C.c_void_p()
        rc = self.lib.vacgpu_batch_create(int(hbm_budget_bytes), int(capacity), C.byref(h))
        if rc != 0:
            raise VacGpuError(f"vacgpu_batch_create failed ({rc})")
        self.h = h
        self.meta = {}
        self._tickets = {}

    def _check(self, rc, what):
        if rc < 0:
            msg = self.lib.vacgpu_batch_last_error(self.h)
            raise VacGpuError(f"{what}: {msg.decode() if msg else rc}")
        return rc

    @staticmethod
    def footprint(nx, physics, gamma=5.0 / 3.0, precision="fp64") -> int:
        lib = load_library()
        cfg = make_config(nx, physics, gamma, precision=precision)
        return int(lib.vacgpu_batch_footprint(C.byref(cfg)))

    def add(self, w_mesh, nx, physics, gamma, dxs, method: MethodSpec, bc="periodic",
            precision="fp64", it=0, t=0.0, itmax=1 << 30, tmax=1e99, tmaxexact=False) -> int:
        cfg = make_config(nx, physics, gamma, bc=bc, precision=precision)
        m = method.to_c(physics)
        w = np.ascontiguousarray(w_mesh, dtype=np.float64)  # the mesh; ghosts are filled on the device
        if w.shape != (nvars(physics, 2), int(nx[1]), int(nx[0])):
            raise ValueError(f"mesh state of shape {w.shape} for a {nx[0]}x{nx[1]} {physics} instance")
        dx0 = np.ascontiguousarray(dxs[0], dtype=np.float64)
        dx1 = np.ascontiguousarray(dxs[1], dtype=np.float64)
        slot = C.c_int()
        rc = self.lib.vacgpu_batch_add_mesh(self.h, C.byref(cfg), C.byref(m), _ptr(dx0), _ptr(dx1),
                                            _ptr(w), it, t, itmax, tmax, int(tmaxexact),
                                            C.byref(slot))
        if rc == -4:
            raise Batch.BudgetFull(self.lib.vacgpu_batch_last_error(self.h).decode())
        self._check(rc, "batch_add")
        self.meta[slot.value] = (tuple(int(n) for n in nx), nvars(physics, 2))
        return slot.value

    def remove(self, slot: int):
        self._check(self.lib.vacgpu_batch_remove(self.h, slot), "batch_remove")
        self.meta.pop(slot, None)

    def step(self, nsteps: int = 1):
        self._check(self.lib.vacgpu_batch_step(self.h, nsteps), "batch_step")

    def sync(self):
        self._check(self.lib.vacgpu_batch_sync(self.h), "batch_sync")

    def status(self, slot: int):
        it = C.c_int()
        t = C.c_double()
        dt = C.c_double()
        done = C.c_int()
        ne = np.zeros(4, dtype=np.int32)
        self._check(self.lib.vacgpu_batch_status(self.h, slot, C.byref(it), C.byref(t), C.byref(dt),
                                                 C.byref(done), ne.ctypes.data_as(_IP)), "status")
        return {"it": it.value, "t": t.value, "dt": dt.value, "done": bool(done.value), "nerror": ne}

    def get_mesh(self, slot: int) -> np.ndarray:
        nx, nw = self.meta[slot]
        w = np.empty((nw, nx[1] + 4, nx[0] + 4))
        self._check(self.lib.vacgpu_batch_get_state(self.h, slot, _ptr(w), None), "get_state")
        return np.ascontiguousarray(w[:, 2:-2, 2:-2])

    def snapshot(self):
        """Enqueue a status snapshot after the steps enqueued so far (no wait)."""
        self._check(self.lib.vacgpu_batch_snapshot(self.h), "batch_snapshot")

    def snapshot_wait(self, wait: bool = True) -> bool:
        """Make the oldest snapshot in flight current; False when none is in
        flight or (wait=False) it has not arrived yet."""
        rc = self.lib.vacgpu_batch_snapshot_wait(self.h, int(wait))
        if rc < 0:
            self._check(rc, "batch_snapshot_wait")
        return rc == 1

    def finished(self) -> list:
        """Slots done in the current snapshot."""
        n = self.lib.vacgpu_batch_finished(self.h, None, 0)
        if n <= 0:
            self._check(n, "batch_finished")
            return []
        out = np.zeros(n, dtype=np.int32)
        n = self.lib.vacgpu_batch_finished(self.h, out.ctypes.data_as(_IP), n)
        return [int(x) for x in out[:n]]

    def fetch_start(self, slot: int) -> int:
        """Enqueue a copy of the instance's mesh after the steps enqueued so
        far; returns a ticket for fetch_ready.  The slot may be removed at once."""
        nx, nw = self.meta[slot]
        tk = C.c_int()
        self._check(self.lib.vacgpu_batch_fetch_start(self.h, slot, 1, C.byref(tk)), "fetch_start")
        self._tickets[tk.value] = np.empty((nw, nx[1], nx[0]))
        return tk.value

    def fetch_ready(self, ticket: int, wait: bool = False):
        """The fetched mesh (nw, ny, nx), or None when (wait=False) it has not arrived."""
        w = self._tickets[ticket]
        rc = self.lib.vacgpu_batch_fetch_ready(self.h, ticket, _ptr(w), None, int(wait))
        if rc < 0:
            self._check(rc, "fetch_ready")
        if rc == 0:
            return None
        del self._tickets[ticket]
        return w

    def graph(self, enable: bool):
        self._check(self.lib.vacgpu_batch_graph(self.h, int(enable)), "batch_graph")

    def tune(self):
        rows = np.zeros(4, dtype=np.int32)
        self._check(self.lib.vacgpu_batch_tune(self.h, rows.ctypes.data_as(_IP)), "batch_tune")
        return [int(r) for r in rows]

    def set_tiles(self, rows):
        r = np.ascontiguousarray(rows, dtype=np.int32)
        self._check(self.lib.vacgpu_batch_set_tiles(self.h, r.ctypes.data_as(_IP)), "set_tiles")

    def set_kernels(self, fused: bool = True, seglen=None):
        """Fused sweeps for eligible instances (default) or predictor +
        corrector tiles for all; seglen: rows per segment per size class."""
        sl = None if seglen is None else np.ascontiguousarray(seglen, dtype=np.int32)
        self._check(self.lib.vacgpu_batch_set_kernels(self.h, int(fused),
                                                      None if sl is None else sl.ctypes.data_as(_IP)),
                    "set_kernels")

    def get_tiles(self):
        rows = np.zeros(4, dtype=np.int32)
        sl = np.zeros(4, dtype=np.int32)
        self._check(self.lib.vacgpu_batch_get_tiles(self.h, rows.ctypes.data_as(_IP),
                                                    sl.ctypes.data_as(_IP)), "get_tiles")
        return [int(r) for r in rows], [int(x) for x in sl]

    def bytes_used(self) -> int:
        return int(self.lib.vacgpu_batch_bytes_used(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.vacgpu_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
