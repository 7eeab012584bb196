"""The Fortran drop-in: VAC's own host with the GPU advect linked in.

Two executables are built from the reference's Fortran sources by
oracle/build_ref.sh: the plain VAC build, and the same host with
vac_amd/fortran/vacgpu_advect.t in place of advect (the CPU advect is renamed
and unused).  Both read the same .ini and the same par file -- the par-file
selectors (typefull, typelimiter, courantpar, divbfix, typeB) reach the GPU
path through VAC's common block -- and their output snapshots must agree
bit for bit, including the t of every step.
"""
import numpy as np
import pytest

from run_ref import have_ref, make_ini, problem_par, run_ref
from vac_amd.initial import orszag_tang, random_fourier

pytestmark = pytest.mark.gpu

NAMES = {6: "rho m1 m2 e b1 b2", 4: "rho m1 m2 e", 8: "rho m1 m2 m3 e b1 b2 b3"}
WALL_X = ["symm", "asymm", "symm", "symm", "asymm", "symm"]
WALL_Y = ["symm", "symm", "asymm", "symm", "symm", "asymm"]

CASES = [
    # (cpu binary, gpu binary, physics, builder, gamma, steps, typefull, limiter, bc, divbfix)
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: orszag_tang(64, g), 5 / 3, 40,
     "tvdlf", "minmod", "periodic", True),
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((48, 40), 61, g), 5 / 3, 25,
     "tvd", "woodward", "periodic", True),
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((40, 32), 62, g), 1.4, 20,
     "tvdlf", "superbee", "cont", False),
    ("vac_hd22_132", "vacgpu_hd22_132", "hd", lambda g: random_fourier((48, 48), 63, g, physics="hd"), 1.4, 30,
     "tvdlf", "woodward", "periodic", True),
    ("vac_mhd33_36", "vacgpu_mhd33_36", "mhd", lambda g: random_fourier((16, 12, 14), 64, g), 5 / 3, 8,
     "tvd", "woodward", "periodic", True),
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((44, 36), 65, g), 5 / 3, 20,
     "tvdmu", "woodward", "periodic", True),
    ("vac_mhd33_36", "vacgpu_mhd33_36", "mhd", lambda g: random_fourier((12, 14, 10), 66, g), 1.4, 6,
     "tvdmu", "minmod", "cont", True),
    ("vac_hd22_132", "vacgpu_hd22_132", "hd", lambda g: random_fourier((44, 40), 71, g, physics="hd"), 1.4, 20,
     "tvd", "woodward", "periodic", True),
    # reflecting walls (symm; the normal momentum and field asymm) and fixed faces
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((40, 36), 68, g), 5 / 3, 15,
     "tvdlf", "minmod", [WALL_X, WALL_X, "fixed", "fixed"], True),
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((36, 44), 69, g), 5 / 3, 12,
     "tvd", "woodward", ["periodic", "periodic", WALL_Y, WALL_Y], True),
    # typelimited='previous': the corrector limits the step's start
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((40, 40), 72, g), 5 / 3, 12,
     "tvdlf", "woodward", "periodic", True, "previous"),
    # an entropy fix on every wave of the Roe dissipation
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((44, 36), 73, g), 5 / 3, 12,
     "tvd", "woodward", "periodic", True, "", "harten"),
    # useprimitive=T: limiting in primitive variables, dp = pR - pL in the Roe solver
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((40, 36), 74, g), 5 / 3, 15,
     "tvdlf", "woodward", "periodic", True, "", "", True),
    ("vac_hd22_132", "vacgpu_hd22_132", "hd", lambda g: random_fourier((36, 40), 75, g, physics="hd"), 1.4, 15,
     "tvdmu", "minmod", "cont", True, "", "", True),
    ("vac_mhd33_36", "vacgpu_mhd33_36", "mhd", lambda g: random_fourier((12, 10, 14), 76, g), 5 / 3, 6,
     "tvdlf", "minmod", "periodic", True, "", "", True),
    # dimsplit=F: one advect over all directions per step
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((40, 36), 77, g), 5 / 3, 15,
     "tvdlf", "woodward", "periodic", True, "", "", False, True),
    ("vac_mhd33_36", "vacgpu_mhd33_36", "mhd", lambda g: random_fourier((12, 10, 14), 78, g), 5 / 3, 6,
     "tvdlf", "minmod", "cont", True, "predictor", "", False, True),
    # first-order extrapolating faces
    ("vac_mhd22_132", "vacgpu_mhd22_132", "mhd", lambda g: random_fourier((36, 40), 79, g), 5 / 3, 12,
     "tvd", "woodward", ["cont1", "fixed1", "grad1", "cont1"], True),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[6]}-{c[7]}-{c[2]}{len(c[3](1.4)[1])}d"
                                             + ("" if isinstance(c[8], str) else "-walls")
                                             + (f"-{c[10]}" if len(c) > 10 and c[10] else "")
                                             + (f"-{c[11]}" if len(c) > 11 and c[11] else "")
                                             + ("-prim" if len(c) > 12 and c[12] else "")
                                             + ("-unsplit" if len(c) > 13 and c[13] else "")
                                             for c in CASES])
def test_vac_with_gpu_advect_matches_vac(case):
    cpu, gpu, physics, build, gamma, steps, typefull, limiter, bc, divbfix = case[:10]
    extra = f" typelimited='{case[10]}'\n" if len(case) > 10 and case[10] else ""
    if len(case) > 11 and case[11]:
        extra += f" typeentropy={len(NAMES[6 if physics == 'mhd' else 4].split())}*'{case[11]}'\n"
    if len(case) > 12 and case[12]:
        extra += " useprimitive=T\n"
    if len(case) > 13 and case[13]:
        extra += " dimsplit=F\n"
    if not (have_ref(cpu) and have_ref(gpu)):
        pytest.skip("reference executables not built (oracle/build_ref.sh)")
    w0, coords = build(gamma)
    nw = w0.shape[0]
    ini = make_ini(w0, coords, physics, gamma, name="dropin")
    typeB = [bc] * (2 * len(coords)) if isinstance(bc, str) else bc
    par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                      typeB=typeB, divbfix=divbfix, every_step=True, extra_method=extra)
    print(f"case {cpu} vs {gpu} {typefull} {limiter} {bc}", flush=True)
    rc = run_ref(cpu, ini, par, timeout=300)
    rg = run_ref(gpu, ini, par, timeout=300)
    tc = np.array([s.t for s in rc.snapshots])
    tg = np.array([s.t for s in rg.snapshots])
    assert np.array_equal(tc, tg), "time sequence differs"
    for sc, sg in zip(rc.snapshots, rg.snapshots):
        assert sc.it == sg.it
        assert np.array_equal(sc.w, sg.w), f"state differs at it={sc.it}"


PROJ_CASES = [
    # (physics builder, gamma, steps, typefull, limiter, bc, divbfix, divbwave, procpar, nproc)
    # the reference's Orszag-Tang recipe (Man/par.html): TVD + woodward,
    # divbfix=F, divbwave=F, nproc=1, procpar=0,0.01,1
    (lambda g: orszag_tang(64, g), 5 / 3, 30, "tvd", "woodward", "periodic", False, False,
     (0.0, 0.01, 1.0), "1"),
    (lambda g: random_fourier((40, 48), 67, g), 1.4, 15, "tvdlf", "minmod", "cont", True, True,
     (0.01, 0.0, 3.0), "1"),
    # as the reference's par/peak22proj: also processed before every save
    # (nproc(4) = 1; this test saves every step), cont faces
    (lambda g: random_fourier((44, 40), 70, g), 5 / 3, 12, "tvd", "woodward", "cont", True, True,
     (-1.0, -1.0, -1.0), "1,0,0,1"),
    # correctors that limit the step's start, which VAC stores before the
    # projection: typelimited='previous', and dimsplit=F ('original' reads as
    # 'previous'; the unsplit corrector's intermediate must not be wold)
    (lambda g: random_fourier((40, 36), 80, g), 5 / 3, 12, "tvdlf", "minmod", "periodic", True,
     True, (0.0, 1e-4, 1.0), "1", " typelimited='previous'\n"),
    (lambda g: random_fourier((36, 40), 81, g), 1.4, 12, "tvdlf", "woodward", "periodic", True,
     True, (0.0, 1e-4, 2.0), "1", " dimsplit=F\n"),
    (lambda g: random_fourier((40, 44), 82, g), 5 / 3, 10, "tvdmu", "superbee", "cont", True,
     True, (0.01, 0.0, 3.0), "1,0,0,1", " typelimited='previous'\n"),
]


@pytest.mark.parametrize("case", PROJ_CASES, ids=[f"{c[3]}-{c[5]}-proj-nproc{c[9]}"
                                                  + ("-" + c[10].split("=")[1].strip().strip("'")
                                                     if len(c) > 10 else "")
                                                  for c in PROJ_CASES])
def test_vac_with_gpu_advect_and_projection_matches_vac(case):
    # VAC built with its Poisson module, plain and with the drop-in, whose
    # projectb projects on the device (vacgpu_process) wherever VAC's process
    # calls it: before a step (process(1) in advect) and before a save.  The
    # CG dot products are summed in another order on the device, so the
    # results agree to rounding: relative L1 <= 1e-12 per variable and the t
    # of every step to 1e-12 relative.
    build, gamma, steps, typefull, limiter, bc, divbfix, divbwave, procpar, nproc = case[:10]
    more = case[10] if len(case) > 10 else ""
    cpu, gpu = "vacp_mhd22_132", "vacgpup_mhd22_132"
    if not (have_ref(cpu) and have_ref(gpu)):
        pytest.skip("reference executables not built (oracle/build_ref.sh ... poisson)")
    w0, coords = build(gamma)
    nw = w0.shape[0]
    ini = make_ini(w0, coords, "mhd", gamma, name="dropin")
    extra = (f" nproc={nproc}\n procpar=" + ",".join(repr(float(v)) for v in procpar) + "\n"
             + more)
    par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter,
                      typeB=[bc] * (2 * len(coords)), divbfix=divbfix, divbwave=divbwave,
                      every_step=True, extra_method=extra)
    print(f"case {cpu} vs {gpu} {typefull} {limiter} {bc} procpar={procpar}", flush=True)
    rc = run_ref(cpu, ini, par, timeout=300)
    rg = run_ref(gpu, ini, par, timeout=300)
    tc = np.array([s.t for s in rc.snapshots])
    tg = np.array([s.t for s in rg.snapshots])
    assert len(tc) == len(tg) == steps + 1
    np.testing.assert_allclose(tg, tc, rtol=1e-12, atol=0)
    for sc, sg in zip(rc.snapshots, rg.snapshots):
        assert sc.it == sg.it
        for iw in range(nw):
            e = np.abs(sg.w[iw] - sc.w[iw]).sum() / np.abs(sc.w[iw]).sum()
            assert e <= 1e-12, f"it={sc.it} variable {iw}: rel L1 {e:.3e}"


def _transfers(out):
    return out.count("vacgpu: state to device"), out.count("vacgpu: state to host")


def test_dropin_keeps_the_state_on_the_device():
    # saves at it = 0 and at the end only: w goes to the device once, at the
    # first sweep, and comes back once, for the final save (src/vac.t:108-111)
    cpu, gpu = "vac_mhd22_132", "vacgpu_mhd22_132"
    if not (have_ref(cpu) and have_ref(gpu)):
        pytest.skip("reference executables not built (oracle/build_ref.sh)")
    w0, coords = random_fourier((48, 44), 83, 5 / 3)
    ini = make_ini(w0, coords, "mhd", 5 / 3, name="dropin")
    steps = 30
    par = problem_par(6, NAMES[6], steps, typefull="tvd", limiter="woodward")
    rc = run_ref(cpu, ini, par, timeout=300)
    rg = run_ref(gpu, ini, par, timeout=300, env_extra={"VACGPU_DROPIN_TRACE": "1"})
    assert [s.it for s in rg.snapshots] == [0, steps]
    for sc, sg in zip(rc.snapshots, rg.snapshots):
        assert sc.t == sg.t
        assert np.array_equal(sc.w, sg.w), f"state differs at it={sc.it}"
    assert _transfers(rg.stdout) == (1, 1), rg.stdout[-2000:]


def test_dropin_fetches_every_step_for_the_residual():
    # residmin > 0: VAC evaluates ||w_new - w_old|| on the host after every
    # step (src/vac.t:90-101), so the drop-in fetches w every step and puts
    # it back at the next; the result stays VAC's bit for bit
    cpu, gpu = "vac_mhd22_132", "vacgpu_mhd22_132"
    if not (have_ref(cpu) and have_ref(gpu)):
        pytest.skip("reference executables not built (oracle/build_ref.sh)")
    w0, coords = random_fourier((40, 36), 84, 1.4)
    ini = make_ini(w0, coords, "mhd", 1.4, name="dropin")
    steps = 10
    par = problem_par(6, NAMES[6], steps, typefull="tvdlf", limiter="minmod")
    par = par.replace(" &stoplist\n", " &stoplist\n residmin=1.D-300\n")
    rc = run_ref(cpu, ini, par, timeout=300)
    rg = run_ref(gpu, ini, par, timeout=300, env_extra={"VACGPU_DROPIN_TRACE": "1"})
    assert [s.it for s in rg.snapshots] == [0, steps]
    assert np.array_equal(rc.snapshots[-1].w, rg.snapshots[-1].w)
    assert rc.snapshots[-1].t == rg.snapshots[-1].t
    assert _transfers(rg.stdout) == (steps, steps), rg.stdout[-2000:]


def test_dropin_stops_on_a_user_source():
    # VAC built with its gravity module (src/vacusr.t.gravity): specialsource
    # adds rho*g to the momentum inside every substep of advect (src/vac.t:592),
    # which the device path does not do, so the run must stop, not run
    # without gravity
    gpu = "vacgpu_mhd22_132_gravity"
    if not have_ref(gpu):
        pytest.skip("reference executables not built (oracle/build_ref.sh with VAC_USR=gravity)")
    w0, coords = random_fourier((32, 28), 85, 5 / 3)
    ini = make_ini(w0, coords, "mhd", 5 / 3, name="dropin",
                   special=(("grav1", 0.0), ("grav2", -1.0)))
    par = problem_par(6, NAMES[6], 5, typefull="tvdlf", limiter="minmod")
    # VAC's die() prints the message and STOPs (src/vacio.t:1108-1116): only
    # the it = 0 save is written
    r = run_ref(gpu, ini, par, timeout=300)
    assert "the linked specialsource changes w" in r.stdout, r.stdout[-2000:]
    assert [s.it for s in r.snapshots] == [0]
    # with g = 0 the module adds nothing: the run proceeds and equals the nul build
    ini0 = make_ini(w0, coords, "mhd", 5 / 3, name="dropin",
                    special=(("grav1", 0.0), ("grav2", 0.0)))
    rg = run_ref(gpu, ini0, par, timeout=300)
    if have_ref("vac_mhd22_132"):
        rc = run_ref("vac_mhd22_132", make_ini(w0, coords, "mhd", 5 / 3, name="dropin"), par,
                     timeout=300)
        assert np.array_equal(rc.snapshots[-1].w, rg.snapshots[-1].w)


W0X = ["symm0", "asymm", "symm0", "symm0", "asymm", "symm0"]
W0Y = ["symm0", "symm0", "asymm", "symm0", "symm0", "asymm"]
MORE_CASES = [
    # (id, cpu, gpu, physics, builder, gamma, steps, typefull, limiter, bc, &methodlist, &paramlist)
    ("tvdlf-symm0-walls", "vac_mhd22_132", "vacgpu_mhd22_132", "mhd",
     lambda g: random_fourier((40, 36), 86, g), 5 / 3, 12, "tvdlf", "minmod",
     [W0X, W0X, "periodic", "periodic"], "", ""),
    ("tvd-symm0-walls", "vac_mhd22_132", "vacgpu_mhd22_132", "mhd",
     lambda g: random_fourier((36, 40), 87, g), 5 / 3, 10, "tvd", "woodward",
     ["periodic", "periodic", W0Y, W0Y], "", ""),
    ("tvdmu-unsplit", "vac_mhd22_132", "vacgpu_mhd22_132", "mhd",
     lambda g: random_fourier((40, 36), 88, g), 5 / 3, 12, "tvdmu", "woodward", "periodic",
     " dimsplit=F\n", ""),
    ("tvdmu-unsplit-3d", "vac_mhd33_36", "vacgpu_mhd33_36", "mhd",
     lambda g: random_fourier((12, 10, 14), 89, g), 5 / 3, 5, "tvdmu", "minmod", "periodic",
     " dimsplit=F\n", ""),
    ("tvdlf-prim-previous", "vac_mhd22_132", "vacgpu_mhd22_132", "mhd",
     lambda g: random_fourier((40, 36), 90, g), 5 / 3, 12, "tvdlf", "woodward", "periodic",
     " useprimitive=T\n typelimited='previous'\n", ""),
    ("tvdmu-prim-predictor-hd", "vac_hd22_132", "vacgpu_hd22_132", "hd",
     lambda g: random_fourier((36, 40), 91, g, physics="hd"), 1.4, 12, "tvdmu", "minmod", "cont",
     " useprimitive=T\n typelimited='predictor'\n", ""),
    ("tvdlf-prim-unsplit-3d", "vac_mhd33_36", "vacgpu_mhd33_36", "mhd",
     lambda g: random_fourier((12, 10, 14), 92, g), 5 / 3, 5, "tvdlf", "minmod", "periodic",
     " useprimitive=T\n dimsplit=F\n", ""),
    # the main loop's time-step controls stay VAC's own with the drop-in
    ("tvd-dtcantgrow-slowsteps", "vac_mhd22_132", "vacgpu_mhd22_132", "mhd",
     lambda g: random_fourier((40, 36), 93, g), 5 / 3, 12, "tvd", "woodward", "periodic",
     "", " dtcantgrow=T\n slowsteps=4\n"),
]


@pytest.mark.parametrize("case", MORE_CASES, ids=[c[0] for c in MORE_CASES])
def test_vac_with_gpu_advect_more_selectors(case):
    (_, cpu, gpu, physics, build, gamma, steps, typefull, limiter, bc, method, param) = case
    if not (have_ref(cpu) and have_ref(gpu)):
        pytest.skip("reference executables not built (oracle/build_ref.sh)")
    w0, coords = build(gamma)
    nw = w0.shape[0]
    ini = make_ini(w0, coords, physics, gamma, name="dropin")
    typeB = [bc] * (2 * len(coords)) if isinstance(bc, str) else bc
    par = problem_par(nw, NAMES[nw], steps, typefull=typefull, limiter=limiter, typeB=typeB,
                      every_step=True, extra_method=method, extra_param=param)
    print(f"case {case[0]}: {cpu} vs {gpu}", flush=True)
    rc = run_ref(cpu, ini, par, timeout=300)
    rg = run_ref(gpu, ini, par, timeout=300)
    assert len(rg.snapshots) == steps + 1, rg.stdout[-2000:]
    tc = np.array([s.t for s in rc.snapshots])
    tg = np.array([s.t for s in rg.snapshots])
    assert np.array_equal(tc, tg), "time sequence differs"
    for sc, sg in zip(rc.snapshots, rg.snapshots):
        assert sc.it == sg.it
        assert np.array_equal(sc.w, sg.w), f"state differs at it={sc.it}"
