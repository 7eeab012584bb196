"""The fused 2D TVDLF sweep kernels against the two-launch form and the oracle.

The fused kernels compute the predictor state of ghost cells at the mesh
cell the boundary rule copies into them, march along y in segments and
exchange face values between lanes; all of that must leave the result
bitwise unchanged.  Cases cover partial strips (nx not a multiple of the 62
cells a wave produces), short and long segments, 'cont' and 'periodic'
faces mixed per face, HD and MHD, both arithmetic tiers, and a grid whose
variables have different boundary types on one face (which must take the
two-launch path and still match the oracle).
"""
import numpy as np
import pytest

import vacref
from vac_amd import vacgpu
from vac_amd.grid import widths_1d
from vac_amd.initial import random_fourier

pytestmark = pytest.mark.gpu

METHOD = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.8)


def run(w0, nx, physics, gamma, dxs, steps, bc, precision="fp64", fused=True, seglen=0,
        method=METHOD):
    g = vacgpu.Grid(nx, physics, gamma, dxs, bc=bc, precision=precision)
    g.set_kernels(fused, seglen)
    g.put_mesh(w0)
    n, t = g.run(method, steps, record_t=True)
    w = g.get_mesh()
    sc = g.scalars()
    g.close()
    return w, t, sc


CASES = [
    # (physics, nx, ny, bc per face or one type, steps, seglen, limiter)
    ("mhd", 61, 37, "periodic", 10, 0, "minmod"),
    ("mhd", 62, 33, "periodic", 10, 4, "woodward"),
    ("mhd", 63, 70, "cont", 8, 5, "superbee"),
    ("mhd", 130, 67, ["periodic", "periodic", "cont", "cont"], 8, 0, "minmod"),
    ("mhd", 125, 9, ["cont", "cont", "periodic", "periodic"], 8, 256, "minmod"),
    ("hd", 200, 41, "periodic", 12, 7, "woodward"),
    ("hd", 7, 130, "cont", 12, 0, "minmod"),
    ("mhd", 5, 5, "periodic", 6, 0, "minmod"),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}{c[1]}x{c[2]}-{c[5]}" for c in CASES])
def test_fused_matches_two_launch_and_oracle(case):
    physics, nx, ny, bc, steps, seglen, limiter = case
    gamma = 5 / 3
    w0, coords = random_fourier((nx, ny), 7 + nx, gamma, physics=physics)
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvdlf", limiter=limiter, courantpar=0.8)
    print(f"case fused {case}", flush=True)
    wf, tf, scf = run(w0, [nx, ny], physics, gamma, dxs, steps, bc, fused=True, seglen=seglen,
                      method=method)
    wu, tu, scu = run(w0, [nx, ny], physics, gamma, dxs, steps, bc, fused=False, method=method)
    assert np.array_equal(tf, tu)
    assert np.array_equal(wf, wu)
    assert np.array_equal(scf["nerror"], scu["nerror"])
    p = vacref.make_params(2, physics, [nx, ny], gamma, typefull="tvdlf", limiter=limiter,
                           bc=bc)
    wr, tr, st = vacref.run(p, w0, dxs, steps)
    assert np.array_equal(tf, tr)
    assert np.array_equal(wf, wr)


@pytest.mark.parametrize("precision", ["fp64", "fp32bulk"])
def test_fused_tiers_match_two_launch(precision):
    # FP64: bitwise.  FP32-bulk: the fused sweep uses its own increment form
    # (FP64 bases plus FP32 increments), so both forms are held to the tier's
    # bar, 1e-5 relative L1 per variable, against each other and the oracle
    w0, coords = random_fourier((190, 150), 3, 1.4)
    dxs = [widths_1d(c) for c in coords]
    print(f"case fused tier {precision}", flush=True)
    wf, tf, _ = run(w0, [190, 150], "mhd", 1.4, dxs, 15, "periodic", precision, fused=True)
    wu, tu, _ = run(w0, [190, 150], "mhd", 1.4, dxs, 15, "periodic", precision, fused=False)
    if precision == "fp64":
        assert np.array_equal(tf, tu)
        assert np.array_equal(wf, wu)
        return
    p = vacref.make_params(2, "mhd", [190, 150], 1.4, typefull="tvdlf", limiter="minmod")
    wr, tr, _ = vacref.run(p, w0, dxs, 15)
    for a, b in ((wf, wu), (wf, wr), (wu, wr)):
        for iw in range(a.shape[0]):
            l1 = np.abs(a[iw] - b[iw]).sum() / np.abs(b[iw]).sum()
            assert l1 <= 1e-5, (iw, l1)
    assert abs(tf[-1] - tr[-1]) <= 1e-5 * tr[-1]


def test_mixed_variable_boundaries_take_two_launch_path():
    # x-low: density periodic, the rest cont (x-high likewise): not eligible
    # for the fused kernels; the result must still be the oracle's
    nx, ny = 40, 30
    w0, coords = random_fourier((nx, ny), 5, 5 / 3, physics="hd")
    dxs = [widths_1d(c) for c in coords]
    mixed = ["periodic", "cont", "cont", "cont"]
    bc = [mixed, mixed, "periodic", "periodic"]
    print("case mixed per-variable boundaries", flush=True)
    w, t, _ = run(w0, [nx, ny], "hd", 5 / 3, dxs, 10, bc, fused=True)
    p = vacref.make_params(2, "hd", [nx, ny], 5 / 3, typefull="tvdlf", limiter="minmod", bc=bc)
    wr, tr, _ = vacref.run(p, w0, dxs, 10)
    assert np.array_equal(t, tr)
    assert np.array_equal(w, wr)


@pytest.mark.parametrize("bc", ["periodic", "cont", ["periodic", "periodic", "cont", "cont"]],
                         ids=["periodic", "cont", "mixed"])
def test_fused_fills_ghost_cells(bc):
    # the fused sweeps write the ghost cells themselves; the whole ixG array
    # (ghosts included) must equal the two-launch path's, which fills them
    # with the separate getboundary kernel
    nx, ny = 70, 45
    w0, coords = random_fourier((nx, ny), 9, 5 / 3)
    dxs = [widths_1d(c) for c in coords]
    out = []
    for fused in (True, False):
        g = vacgpu.Grid([nx, ny], "mhd", 5 / 3, dxs, bc=bc)
        g.set_kernels(fused, 5)
        g.put_mesh(w0)
        g.run(METHOD, 7)
        out.append(g.get_state())
        g.close()
    print(f"case fused ghosts {bc}", flush=True)
    assert np.array_equal(out[0], out[1])


TVD_CASES = [
    # (nx, ny, bc, limiter, precision, steps)
    (61, 37, "periodic", "woodward", "fp64", 9),
    (64, 40, "cont", "superbee", "fp64", 9),
    (130, 67, ["periodic", "periodic", "cont", "cont"], "minmod", "fp64", 6),
    (45, 90, "periodic", "woodward", "fp32bulk", 8),
]


@pytest.mark.parametrize("case", TVD_CASES, ids=[f"{c[0]}x{c[1]}-{c[3]}-{c[4]}" for c in TVD_CASES])
def test_tvd_folded_sources_match(case):
    # the TVD sweeps fold the first-sweep half source and the first half of
    # the Ryu step; the whole ixG array must equal the separate-kernel path,
    # and the FP64 result the oracle's
    nx, ny, bc, limiter, precision, steps = case
    w0, coords = random_fourier((nx, ny), 21 + nx, 5 / 3)
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvd", limiter=limiter, courantpar=0.8)
    print(f"case tvd fold {case}", flush=True)
    out = []
    for fused in (True, False):
        g = vacgpu.Grid([nx, ny], "mhd", 5 / 3, dxs, bc=bc, precision=precision)
        g.set_kernels(fused)
        g.put_mesh(w0)
        n, t = g.run(method, steps, record_t=True)
        out.append((g.get_state(), t))
        g.close()
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    if precision == "fp64":
        p = vacref.make_params(2, "mhd", [nx, ny], 5 / 3, typefull="tvd", limiter=limiter, bc=bc)
        wr, tr, _ = vacref.run(p, w0, dxs, steps)
        assert np.array_equal(out[0][1], tr)
        assert np.array_equal(out[0][0][:, 2:-2, 2:-2], wr)


GHOST_CASES = [
    # (physics, shape, bc, limiter, steps): TVDLF grids the fused 2D kernels
    # do not take, swept by the two-launch kernels that write ghosts themselves
    ("mhd", (20, 14, 11), "periodic", "minmod", 5),
    ("mhd", (13, 18, 9), ["cont", "cont", "periodic", "periodic", "cont", "cont"], "woodward", 5),
    ("hd", (24, 10, 19), "cont", "superbee", 6),
    ("hd", (33, 21, 16), "periodic", "minmod", 6),
]


@pytest.mark.parametrize("case", GHOST_CASES,
                         ids=[f"{c[0]}{'x'.join(map(str, c[1]))}-{c[3]}" for c in GHOST_CASES])
def test_two_launch_tvdlf_writes_ghost_cells(case):
    # with one boundary type per face the predictor and corrector kernels
    # write the ghost cells of their outputs; the whole ixG array must equal
    # the path that fills them with the getboundary kernel, and the mesh the
    # oracle's
    physics, shape, bc, limiter, steps = case
    w0, coords = random_fourier(shape, 31 + shape[0], 5 / 3, physics=physics)
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvdlf", limiter=limiter, courantpar=0.8)
    print(f"case two-launch ghosts {case}", flush=True)
    out = []
    for fused in (True, False):
        g = vacgpu.Grid(list(shape), physics, 5 / 3, dxs, bc=bc)
        g.set_kernels(fused)
        g.put_mesh(w0)
        n, t = g.run(method, steps, record_t=True)
        out.append((g.get_state(), t))
        g.close()
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
    p = vacref.make_params(len(shape), physics, list(shape), 5 / 3, typefull="tvdlf",
                           limiter=limiter, bc=bc)
    wr, tr, _ = vacref.run(p, w0, dxs, steps)
    assert np.array_equal(out[0][1], tr)
    mesh = (slice(None),) + (slice(2, -2),) * len(shape)
    assert np.array_equal(out[0][0][mesh], wr)


@pytest.mark.parametrize("bc", ["periodic", "cont"])
def test_tvd_3d_ghost_writes_and_ryu_pass(bc):
    # 3D TVD: the x sweep writes its ghost cells and the last (z) sweep's Ryu
    # step runs as one pass over wold and wnew; the whole ixG array must equal
    # the path with separate boundary passes and Ryu halves
    shape = (18, 12, 10)
    w0, coords = random_fourier(shape, 41, 5 / 3)
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvd", limiter="woodward", courantpar=0.8)
    print(f"case tvd 3d ghosts {bc}", flush=True)
    out = []
    for fused in (True, False):
        g = vacgpu.Grid(list(shape), "mhd", 5 / 3, dxs, bc=bc)
        g.set_kernels(fused)
        g.put_mesh(w0)
        n, t = g.run(method, 4, record_t=True)
        out.append((g.get_state(), t))
        g.close()
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][0], out[1][0])
