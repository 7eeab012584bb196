"""ctypes binding of the C restatement in oracle/vacref.c.

Test infrastructure only (tests, the smoke check and bench.py's CPU-baseline
leg use it as the checker).  Arrays follow VAC's layout with the leading
dimensions equal to the grid: w[nw, (nz+4,) ny+4, nx+4], x fastest.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libvacref.so")

HD, MHD = 0, 1
TVDLF, TVD, TVDMU = 0, 1, 2
LIMITERS = {"minmod": 0, "woodward": 1, "superbee": 2}
CONT, PERIODIC, SYMM, ASYMM, FIXED, CONT1, FIXED1, GRAD1, SYMM0 = 0, 1, 2, 3, 4, 5, 6, 7, 8
BC_CODES = {"cont": CONT, "periodic": PERIODIC, "symm": SYMM, "asymm": ASYMM, "fixed": FIXED,
            "cont1": CONT1, "fixed1": FIXED1, "grad1": GRAD1, "symm0": SYMM0}


class Params(C.Structure):
    _fields_ = [
        ("ndim", C.c_int), ("phys", C.c_int), ("nx", C.c_int * 3),
        ("gamma", C.c_double), ("typefull", C.c_int), ("limiter", C.c_int * 8),
        ("bc", (C.c_int * 8) * 6), ("divbfix", C.c_int), ("divbwave", C.c_int),
        ("sourceunsplit", C.c_int), ("dimsplit_xy", C.c_int),
        ("courantpar", C.c_double), ("smallpcoeff", C.c_double),
        ("dtpar", C.c_double), ("tmax", C.c_double), ("tmaxexact", C.c_int),
        ("itmax", C.c_int), ("nproc1", C.c_int), ("procpar", C.c_double * 3),
        ("typelimited", C.c_int), ("typeentropy", C.c_int * 8), ("entropycoef", C.c_double * 8),
        ("useprimitive", C.c_int), ("unsplit", C.c_int),
        ("dtcantgrow", C.c_int), ("slowsteps", C.c_int),
    ]


class State(C.Structure):
    _fields_ = [
        ("it", C.c_int), ("itmin", C.c_int), ("t", C.c_double), ("dt", C.c_double),
        ("smallp", C.c_double), ("dtcourant", C.c_double * 3),
        ("nerror", C.c_int * 4), ("divbmax", C.c_double), ("matvecmax", C.c_int),
        ("typestop", C.c_int), ("proj_info", C.c_int), ("proj_iter", C.c_int),
        ("proj_resid", C.c_double), ("nproj", C.c_int),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE, "libvacref.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.vacref_run.argtypes = [C.POINTER(Params), C.POINTER(State), dp, dp, dp, dp,
                                 C.c_int, dp]
        L.vacref_run.restype = C.c_int
        L.vacref_getboundary.argtypes = [C.POINTER(Params), dp]
        L.vacref_getboundary.restype = C.c_int
        L.vacref_projectb.argtypes = [C.POINTER(Params), C.POINTER(State), dp, dp, dp, dp]
        L.vacref_projectb.restype = C.c_int
        L.vacref_getdt.argtypes = [C.POINTER(Params), C.POINTER(State), dp, dp, dp, dp]
        L.vacref_getdt.restype = C.c_int
        L.vacref_limit.argtypes = [C.c_int, C.c_double, C.c_double]
        L.vacref_limit.restype = C.c_double
        L.vacref_roe_identity.argtypes = [C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                          dp, dp, dp, dp]
        L.vacref_roe_identity.restype = C.c_int
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(ndim, physics, nx, gamma, typefull="tvdlf", limiter="minmod",
                bc="periodic", divbfix=True, divbwave=True, sourceunsplit=True,
                dimsplit_xy=False, courantpar=0.8, smallpcoeff=1e-6, dtpar=-1.0,
                tmax=1e99, tmaxexact=False, itmax=1000000, nproc1=0,
                procpar=(-1.0, -1.0, -1.0), typelimited="original", typeentropy="nul",
                entropycoef=None, useprimitive=False, unsplit=False, dtcantgrow=False,
                slowsteps=0) -> Params:
    p = Params()
    p.ndim = ndim
    p.phys = MHD if physics == "mhd" else HD
    for d in range(ndim):
        p.nx[d] = int(nx[d])
    p.gamma = gamma
    p.typefull = {"tvdlf": TVDLF, "tvd": TVD, "tvdmu": TVDMU}[typefull]
    lims = [limiter] * 8 if isinstance(limiter, str) else list(limiter) + ["minmod"] * 8
    for i in range(8):
        p.limiter[i] = LIMITERS[lims[i]]
    # bc: one type for all faces, or one entry per face; an entry is a type
    # or a list of types per variable
    bcs = [bc] * 6 if isinstance(bc, str) else list(bc)
    for f in range(6):
        spec = bcs[f] if f < len(bcs) else "cont"
        for iw in range(8):
            name = spec if isinstance(spec, str) else spec[min(iw, len(spec) - 1)]
            p.bc[f][iw] = BC_CODES[name]
    p.divbfix = int(divbfix)
    p.divbwave = int(divbwave)
    p.sourceunsplit = int(sourceunsplit)
    p.dimsplit_xy = int(dimsplit_xy)
    p.courantpar = courantpar
    p.smallpcoeff = smallpcoeff
    p.dtpar = dtpar
    p.tmax = tmax
    p.tmaxexact = int(tmaxexact)
    p.itmax = itmax
    p.typelimited = {"original": 0, "previous": 1, "predictor": 2}[typelimited]
    ents = [typeentropy] * 8 if isinstance(typeentropy, str) else list(typeentropy) + ["nul"] * 8
    for i in range(8):
        p.typeentropy[i] = ENTROPY[ents[i]]
    coef = entropy_defaults(physics, ndim) if entropycoef is None else list(entropycoef) + [-1.0] * 8
    for i in range(8):
        p.entropycoef[i] = float(coef[i])
    p.useprimitive = int(useprimitive)
    p.unsplit = int(unsplit)
    p.dtcantgrow = int(dtcantgrow)
    p.slowsteps = int(slowsteps)
    p.nproc1 = int(nproc1)
    for i in range(3):
        p.procpar[i] = float(procpar[i])
    return p


ENTROPY = {"nul": 0, "harten": 1, "powell": 2, "ratio": 3, "yee": 4}


def entropy_defaults(physics, ndim):
    """physini's entropycoef (src/vacphys.mhd0.t:18-29, src/vacphys.t.hd:20-28):
    0.2 for the fast and slow (sound) waves, 0.4 for the Alfven waves, -1 else.
    The literals there are default (single precision) reals, so the stored
    doubles are 0.2 and 0.4 rounded to single precision."""
    f2, f4 = float(np.float32(0.2)), float(np.float32(0.4))
    if physics == "mhd":
        c = [f2, f2, f2, f2, -1.0, -1.0, f4, f4]
    else:
        c = [f2, f2] + [-1.0] * 6
    return c


def bc_spec(v):
    """The boundary spec a fixture stores: one type name, or a JSON list with
    one entry per face (a type, or a list of types per variable)."""
    import json
    s = str(v)
    return json.loads(s) if s[:1] in ("[", '"') else s


def pad_state(w_mesh: np.ndarray) -> np.ndarray:
    """Embed a mesh state (nw, ..., nx) into a ghost-padded ixG array."""
    ndim = w_mesh.ndim - 1
    pad = [(0, 0)] + [(2, 2)] * ndim
    return np.ascontiguousarray(np.pad(w_mesh, pad))


def mesh_of(w_full: np.ndarray) -> np.ndarray:
    ndim = w_full.ndim - 1
    sl = (slice(None),) + (slice(2, -2),) * ndim
    return np.ascontiguousarray(w_full[sl])


def run(params: Params, w_mesh: np.ndarray, dxs: list, nsteps: int, it0=0, t0=0.0,
        smallp=-1.0):
    """Run the restated main loop; returns (w_mesh_out, t_hist, state)."""
    L = lib()
    w = pad_state(np.asarray(w_mesh, dtype=np.float64))
    st = State()
    st.it = it0
    st.itmin = it0
    st.t = t0
    st.dt = 0.0
    st.smallp = smallp
    for d in range(3):
        st.dtcourant[d] = 1e99
    dx = [np.ascontiguousarray(d_, dtype=np.float64) for d_ in dxs]
    while len(dx) < 3:
        dx.append(np.ones(1))
    hist = np.zeros(nsteps + 1)
    n = L.vacref_run(C.byref(params), C.byref(st), _ptr(w), _ptr(dx[0]), _ptr(dx[1]),
                     _ptr(dx[2]), nsteps, _ptr(hist))
    if n < 0:
        raise RuntimeError(f"vacref_run failed ({n})")
    return mesh_of(w), hist[: n + 1], st


def projectb(params: Params, w_mesh: np.ndarray, dxs: list, state: State | None = None):
    """One projectb call on a mesh state (ghosts filled as getboundary does);
    returns (w_full_out, state) with the whole ixG array."""
    L = lib()
    w = pad_state(np.asarray(w_mesh, dtype=np.float64))
    st = state if state is not None else State()
    dx = [np.ascontiguousarray(d_, dtype=np.float64) for d_ in dxs]
    while len(dx) < 3:
        dx.append(np.ones(1))
    # fill the ghost cells first, as the state of a running VAC has them
    L.vacref_getboundary(C.byref(params), _ptr(w))
    rc = L.vacref_projectb(C.byref(params), C.byref(st), _ptr(w), _ptr(dx[0]), _ptr(dx[1]),
                           _ptr(dx[2]))
    if rc:
        raise RuntimeError(f"vacref_projectb failed ({rc})")
    return w, st


def limit(kind: str, a: float, b: float) -> float:
    return lib().vacref_limit(LIMITERS[kind], a, b)


def roe_identity(ndim, gamma, idim, wL, wR, divbwave=True, smallp=0.0):
    nw = 6 if ndim == 2 else 8
    wl = np.ascontiguousarray(wL, dtype=np.float64)
    wr = np.ascontiguousarray(wR, dtype=np.float64)
    out = np.zeros(nw)
    a = np.zeros(nw)
    rc = lib().vacref_roe_identity(ndim, gamma, idim, int(divbwave), smallp, _ptr(wl),
                                   _ptr(wr), _ptr(out), _ptr(a))
    if rc:
        raise RuntimeError("roe identity failed")
    return out, a
