"""Benchmark of the MI355X advect path on its headline workload.

Workload: 2D ideal MHD on a 2048x2048 periodic grid, FP64, gamma=5/3, TVD
(Roe-type) with the woodward limiter, courantpar=0.8, 200 steps, from the
seeded random-Fourier initial state (rho = 1 + 0.2*sum of 8 modes, p = 1,
v and B = 0.1*sum of 8 modes each, B from a vector potential).

A step is one full VAC time step (Courant dt, then one dimensionally split
sweep per direction, each with the Roe dissipation and the Powell / Ryu
source terms).  Metric: Mcell-updates/s = nx*ny*steps / seconds / 1e6.

    python bench.py [--gpus 1] [--steps 200] [--warmup 5] [--no-cpu] [--precision fp64]
                    [--no-tvdlf] [--no-tvdmu] [--no-unsplit] [--no-projection] [--no-3d]
                    [--no-snapshots] [--no-batched] [--no-serving] [--no-dropin]

Prints one JSON line.  Beside the headline it reports the batched workload
(512 independent 2D MHD grids, sides log-uniform in [64, 512], TVDLF +
minmod, 100 steps each, FP64 and FP32-bulk) and the served workload (open-loop
Poisson arrivals, 2000 requests at each of 50, 200, 1000 and 4000 requests/s,
of 2D HD or MHD solves, side log-uniform in [32, 1024], end time uniform in
[0.05, 0.5], FP64 or FP32-bulk) with p50/p99 latency and the offered load.  The state is resident in HBM before the timed region
starts; the timed region is the device-side time loop (vacgpu_run) between
two stream synchronisations.  The TVD sweep kernel's average duration is
measured with HIP events on the library's own stream in a separate pass.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX = NY = 2048
STEPS = 200
SEED = 20251203
GAMMA = 5.0 / 3.0
NW = 6
# algorithmic bytes per cell of the TVD sweep launches, in VAC's dataflow
# (every substep reads its inputs and writes its output once, FP64):
# centdiff (read wCT, read w, write w) + tvdlimit (read wCT, read w, write w)
# = 6*nw*8 per launch; the first sweep of a step also does the half-source
# step (read w, write wCT: 2*nw*8) and an x sweep that ends a step (every
# other step) the first half of the Ryu step (read wold, read w, write wt:
# 3*nw*8), so per launch on average 6*nw*8 + 2*nw*8/2 + 3*nw*8/4
B_ALG_TVD_SWEEP = 6 * NW * 8 + (2 * NW * 8) // 2 + (3 * NW * 8) // 4
# a whole TVD step (every array pass counted: wold copy 2nw, half source 2nw,
# 6nw per sweep, Ryu 6nw) = (6*ndim + 10)*nw*8
B_ALG_TVD_STEP = (6 * 2 + 10) * NW * 8
HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed counter runs
    (FETCH_SIZE and WRITE_SIZE, calibrated on a 1 GiB copy), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            return json.load(f)[kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


# FP64 VALU issue peak of MI355X: 78.6 TFLOP/s of 64-lane FMA wave
# instructions (2 flops per lane) = 614 G wave-instructions/s; a non-packed
# VALU instruction of any kind takes the same 4 cycles of a 16-lane SIMD
VALU_PEAK_GINST = 78.6e3 / 128
VALU_FILE = os.path.join(ROOT, "profiles", "valu.json")
SERVE_REQUESTS = 2000  # requests per arrival rate of the served workload


def valu_roofline(kernel, kernel_ms):
    """The dominant kernel against the VALU issue ceiling: its VALU
    wave-instructions per launch (SQ_INSTS_VALU from the committed counter
    run, profiles/valu.json) over its live average duration."""
    try:
        with open(VALU_FILE) as f:
            insts = json.load(f)[kernel]["valu_insts_per_launch"]
    except (OSError, KeyError, ValueError):
        return None
    if not kernel_ms:
        return None
    rate = insts / (kernel_ms * 1e-3) / 1e9
    return {"bound": "fp64-valu", "achieved": round(rate, 1), "peak": round(VALU_PEAK_GINST, 1),
            "unit": "G wave-instructions/s", "frac": round(rate / VALU_PEAK_GINST, 4),
            "valu_insts_per_launch": insts, "kernel": kernel,
            "source": "profiles/valu.json (rocprofv3 --pmc SQ_INSTS_VALU of bench.py --probe)"}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores() -> list:
    """The cores this process may use, at most 16 (one GPU's share of the host)."""
    cores = sorted(os.sched_getaffinity(0))
    return cores[:16]


def _ref_runs(binary, ini, par, cores):
    """One reference run per core in `cores`, all at once, each pinned to its
    core; returns their advance times (s)."""
    from concurrent.futures import ThreadPoolExecutor

    from run_ref import run_ref

    with ThreadPoolExecutor(max_workers=len(cores)) as ex:
        futs = [ex.submit(run_ref, binary, ini, par, 900, None, None, c) for c in cores]
        return [f.result().advance_seconds for f in futs]


def cpu_baseline(w0, coords):
    """The reference VAC (compiled -O3 for x86-64-v3 by oracle/build_ref.sh) on
    the host's cores: the headline workload (2048^2 TVD + woodward) for one
    step, on one core and as N independent processes pinned one per core
    (aggregate throughput = N x cells / the slowest advance time); the same
    for config 1 (Orszag-Tang 128^2, TVDLF + minmod, 100 steps).  VAC is
    serial (no threads), so N processes is its multi-core form.  Advance time
    only ("Finish Advance" - "Start Advance", src/vac.t:52,115).  Falls back to
    the C restatement on one core if the executables are missing.  Config 3
    (3D TVDLF) is timed the same way on a 128^3 sample of its state."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from run_ref import have_ref, make_ini, problem_par
    from vac_amd.initial import orszag_tang

    cells = NX * NY
    if have_ref("vacO3_mhd22_2052") and have_ref("vacO3_mhd22_132"):
        cores = host_cores()
        n = len(cores)
        ini = make_ini(w0, coords, "mhd", GAMMA, name="bench")
        par = problem_par(6, "rho m1 m2 e b1 b2", 1, typefull="tvd", limiter="woodward",
                          courantpar=0.8)
        one = _ref_runs("vacO3_mhd22_2052", ini, par, cores[:1])[0]
        many = _ref_runs("vacO3_mhd22_2052", ini, par, cores)
        w1, c1 = orszag_tang(128, GAMMA)
        ini1 = make_ini(w1, c1, "mhd", GAMMA, name="bench")
        par1 = problem_par(6, "rho m1 m2 e b1 b2", 100, typefull="tvdlf", limiter="minmod",
                           courantpar=0.8)
        one1 = _ref_runs("vacO3_mhd22_132", ini1, par1, cores[:1])[0]
        many1 = _ref_runs("vacO3_mhd22_132", ini1, par1, cores)
        cfg3 = None
        if have_ref("vacO3_mhd33_132"):
            # config 3 (3D MHD TVDLF + minmod, courantpar 0.6) on a 128^3 sample
            # of its random-Fourier state: 2 steps per process
            from vac_amd.initial import random_fourier

            w3, c3 = random_fourier((128, 128, 128), SEED + 3, GAMMA)
            ini3 = make_ini(w3, c3, "mhd", GAMMA, name="bench")
            par3 = problem_par(8, "rho m1 m2 m3 e b1 b2 b3", 2, typefull="tvdlf",
                               limiter="minmod", courantpar=0.6, typeB=["periodic"] * 6)
            one3 = _ref_runs("vacO3_mhd33_132", ini3, par3, cores[:1])[0]
            many3 = _ref_runs("vacO3_mhd33_132", ini3, par3, cores)
            cfg3 = {"single_core": round(128 ** 3 * 2 / one3 / 1e6, 3),
                    "aggregate": round(n * 128 ** 3 * 2 / max(many3) / 1e6, 3),
                    "cores": n, "unit": "Mcell-updates/s",
                    "sample": "128^3 of config 3's random-Fourier state, 2 steps per process"}
        return {"value": round(n * cells / max(many) / 1e6, 4), "unit": "Mcell-updates/s",
                "cores": n, "kind": "reference",
                "sample": (f"VAC 4.52 (flang -O3 -march=x86-64-v3), {n} independent processes "
                           "pinned one per core, each one step of the 2048x2048 TVD+woodward "
                           "workload (of 200); advance time only"),
                "cpu_model": cpu_model(),
                "single_core": {"value": round(cells / one / 1e6, 4), "seconds_per_step": one},
                "per_process_seconds": [round(v, 3) for v in many],
                "config1_ot128_tvdlf_100": {
                    "single_core": round(128 * 128 * 100 / one1 / 1e6, 3),
                    "aggregate": round(n * 128 * 128 * 100 / max(many1) / 1e6, 3),
                    "cores": n, "unit": "Mcell-updates/s"},
                "config3_mhd3d_tvdlf": cfg3}
    import vacref
    from vac_amd.grid import widths_1d

    m = 512
    p = vacref.make_params(2, "mhd", [m, m], GAMMA, typefull="tvd", limiter="woodward")
    sub = np.ascontiguousarray(w0[:, :m, :m])
    t0 = time.perf_counter()
    vacref.run(p, sub, [widths_1d(c[:m]) for c in coords], 2)
    dt = time.perf_counter() - t0
    return {"value": m * m * 2 / dt / 1e6, "unit": "Mcell-updates/s", "cores": 1,
            "kind": "port", "sample": "C restatement, 512x512 TVD+woodward, 2 steps",
            "cpu_model": cpu_model()}


def dropin_2048(w0, coords, steps):
    """VAC's own host loop (the reference's Fortran main program, compiled
    here) with the drop-in advect on the GPU, on the headline workload.  The
    drop-in's trace prints the wall clock at the start of every step; the
    per-step figure is the mean interval between those lines from the second
    step on, so the first step's host-side getdt_courant, the library load
    and the state's upload are left out.  The transfers are counted from the
    same trace (one upload, one fetch for the final save).  VAC's loop itself
    copies w into wold on the host every step (src/vac.t:85)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from run_ref import have_ref, make_ini, problem_par, run_ref

    name = "vacgpu_mhd22_2052"
    if not have_ref(name):
        return {"skipped": f"oracle/_ref/{name} not built"}
    ini = make_ini(w0, coords, "mhd", GAMMA, name="bench")
    par = problem_par(6, "rho m1 m2 e b1 b2", 1 + steps, typefull="tvd", limiter="woodward",
                      courantpar=0.8)
    r = run_ref(name, ini, par, timeout=900, env_extra={"VACGPU_DROPIN_TRACE": "1"})
    wall = [float(line.split("wall_s=")[1]) for line in r.stdout.splitlines()
            if line.startswith("vacgpu: step at it=")]
    if len(wall) < 3:
        return {"error": f"{len(wall)} step lines in the drop-in trace"}
    per = (wall[-1] - wall[1]) / (len(wall) - 2)
    return {"mcell_updates_per_s": round(NX * NY / per / 1e6, 2),
            "ms_per_step": round(per * 1e3, 3),
            "first_step_s": round(wall[1] - wall[0], 3),
            "steps_timed": len(wall) - 2,
            "advance_s": round(r.advance_seconds, 3),
            "state_transfers_to_device_to_host": [r.stdout.count("vacgpu: state to device"),
                                                  r.stdout.count("vacgpu: state to host")],
            "workload": (f"VAC 4.52 host loop + GPU advect (vacgpu_advect.t), 2048x2048 TVD+"
                         f"woodward, {1 + steps} steps (the {len(wall) - 2} after the first "
                         f"timed from the trace's per-step wall clock), saves at the start "
                         f"and end")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=STEPS)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32bulk"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe", action="store_true",
                    help="run only a short workload for hardware-counter passes: a 1 GiB "
                         "device copy (calibration), then 3 steps each of the 2048^2 TVD "
                         "and TVDLF workloads; prints no JSON")
    ap.add_argument("--probe-3d", action="store_true",
                    help="as --probe, with 3 steps of the 256^3 TVDLF workload instead of the "
                         "2048^2 ones")
    ap.add_argument("--no-tvdlf", action="store_true")
    ap.add_argument("--no-tvdmu", action="store_true")
    ap.add_argument("--no-unsplit", action="store_true")
    ap.add_argument("--no-projection", action="store_true")
    ap.add_argument("--no-3d", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-snapshots", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(independent grids shard by instance)", file=sys.stderr)
        sys.exit(2)

    from vac_amd import vacgpu
    from vac_amd.grid import widths_1d
    from vac_amd.initial import random_fourier

    if args.probe_3d:
        probe_3d(args.precision)
        return
    w0, coords = random_fourier((NX, NY), SEED, GAMMA)
    if args.probe:
        probe(w0, [widths_1d(c) for c in coords], args.precision)
        return
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvd", limiter="woodward", courantpar=0.8)
    grid = vacgpu.Grid([NX, NY], "mhd", GAMMA, dxs, precision=args.precision)

    # warm-up (first launches, clocks), then reset to the initial state
    grid.put_mesh(w0)
    if args.warmup > 0:
        grid.run(method, args.warmup)
    grid.close()
    grid = vacgpu.Grid([NX, NY], "mhd", GAMMA, dxs, precision=args.precision)
    grid.put_mesh(w0)

    t0 = time.perf_counter()
    n, _ = grid.run(method, args.steps)  # returns after a stream synchronisation
    elapsed = time.perf_counter() - t0
    if n != args.steps:
        raise RuntimeError(f"ran {n} of {args.steps} steps")
    sc = grid.scalars()
    it, t_end = grid.get_time()
    ms_per_step = elapsed / args.steps * 1e3
    value = NX * NY * args.steps / elapsed / 1e6

    # dominant-kernel timing with HIP events on the library stream, over the
    # same steps from the same state as the timed run (the sweep kernels'
    # duration depends on the state: the first steps of this one are slower)
    grid.put_mesh(w0)
    grid.set_time(0, 0.0)
    grid.profile(True)
    grid.run(method, args.steps)
    k_ms, k_cnt = grid.kernel_time("tvd_sweep")
    other = {k: grid.kernel_time(k) for k in ("ryu", "bound", "reduce")}
    grid.profile(False)
    grid.close()
    achieved = B_ALG_TVD_SWEEP * NX * NY / (k_ms * 1e-3) / 1e9 if k_ms > 0 else 0.0

    out = {
        "metric": "Mcell-updates/s per step (2D MHD, TVD Roe + woodward, 2048^2)",
        "value": round(value, 3),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-bulk/f64-accumulate",
        "data": "synthetic (seeded random-Fourier MHD state)",
        "config": {"workload": "2D ideal MHD 2048x2048 periodic, TVD (Roe) + woodward, "
                               f"courantpar 0.8, {args.steps} steps, gamma 5/3",
                   "nx": NX, "ny": NY, "seed": SEED, "precision": args.precision},
        "t_end": t_end,
        "nerror": [int(v) for v in sc["nerror"]],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic("k_tvd_sweep"),
                     "traffic_source": "profiles/traffic.json (rocprofv3 --pmc FETCH_SIZE, "
                                       "WRITE_SIZE of bench.py --probe), not this run",
                     "kernel": "k_tvd_sweep",
                     "kernel_avg_ms": round(k_ms, 5), "kernel_launches": k_cnt,
                     "b_alg_per_cell": B_ALG_TVD_SWEEP},
        "roofline_valu": valu_roofline("k_tvd_sweep", k_ms),
        "other_kernels_avg_ms": {k: round(v[0], 5) for k, v in other.items()},
        "step_effective_hbm_frac": round(B_ALG_TVD_STEP * NX * NY / (ms_per_step * 1e-3) / 1e9
                                         / HBM_PEAK_GBS, 4),
    }
    if not args.no_tvdlf:
        out["tvdlf_2048"] = tvdlf_2048(w0, dxs, min(args.steps, 100))
    if not args.no_tvdmu:
        out["tvdmu_2048"] = tvdmu_2048(w0, dxs, min(args.steps, 50))
    if not args.no_unsplit:
        out["tvdlf_unsplit_2048"] = tvdlf_unsplit_2048(w0, dxs, min(args.steps, 50))
    if not args.no_projection:
        out["ot_projection_2048"] = ot_projection_2048()
    if not args.no_3d:
        out["mhd3d_256"] = mhd3d_256()
    if not args.no_snapshots:
        out["snapshots"] = snapshot_overhead(w0, coords, dxs, method)
    if not args.no_batched:
        out["batched"] = batched()
    if not args.no_serving:
        out["serving"] = serving()
    if not args.no_dropin:
        out["dropin_2048"] = dropin_2048(w0, coords, max(50, min(args.steps, 200)))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w0, coords)
    print(json.dumps(out))


def probe(w0, dxs, precision):
    """Short run for counter passes (rocprofv3 --pmc): the copy kernel moves
    exactly 1 GiB in and 1 GiB out, which calibrates FETCH_SIZE and
    WRITE_SIZE; then the two 2048^2 workloads for 3 steps each."""
    import torch

    from vac_amd import vacgpu

    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    for typefull, lim in (("tvd", "woodward"), ("tvdlf", "minmod")):
        g = vacgpu.Grid([NX, NY], "mhd", GAMMA, dxs, precision=precision)
        g.put_mesh(w0)
        g.run(vacgpu.MethodSpec(typefull=typefull, limiter=lim, courantpar=0.8), 3)
        g.close()
    print("probe done", flush=True)


def probe_3d(precision):
    """Short run for counter passes on config 3: the 1 GiB copy (calibration),
    then 3 steps of the 256^3 TVDLF + minmod workload."""
    import torch

    from vac_amd import vacgpu
    from vac_amd.grid import widths_1d
    from vac_amd.initial import random_fourier

    a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    n = 256
    w0, coords = random_fourier((n, n, n), SEED + 3, GAMMA)
    g = vacgpu.Grid([n, n, n], "mhd", GAMMA, [widths_1d(c) for c in coords], precision=precision)
    g.put_mesh(w0)
    g.run(vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.6), 3)
    g.close()
    print("probe-3d done", flush=True)


def errors(a, ref):
    """Per conserved variable: relative L1 = sum|a-ref|/sum|ref|, Linf = max|a-ref|."""
    nw = ref.shape[0]
    d = np.abs(a.reshape(nw, -1) - ref.reshape(nw, -1))
    l1 = d.sum(axis=1) / np.abs(ref.reshape(nw, -1)).sum(axis=1)
    return {"rel_l1": [float(f"{v:.3e}") for v in l1],
            "linf": [float(f"{v:.3e}") for v in d.max(axis=1)]}


def timed_run(shape, physics, dxs, w0, method, precision, steps, kernels):
    """Warm up, reset, time `steps` full steps; then time the listed kernels."""
    from vac_amd import vacgpu

    g = vacgpu.Grid(shape, physics, GAMMA, dxs, precision=precision)
    g.put_mesh(w0)
    g.run(method, 3)
    g.put_mesh(w0)
    g.set_time(0, 0.0)
    t0 = time.perf_counter()
    n, _ = g.run(method, steps)
    elapsed = time.perf_counter() - t0
    if n != steps:
        raise RuntimeError(f"ran {n} of {steps} steps")
    w = g.get_mesh()
    g.put_mesh(w0)
    g.set_time(0, 0.0)
    g.profile(True)
    g.run(method, steps)
    kt = {k: round(g.kernel_time(k)[0], 5) for k in kernels}
    g.profile(False)
    g.close()
    return elapsed, w, kt


def tvdlf_2048(w0, dxs, steps):
    """2D MHD 2048^2, TVDLF + minmod, FP64 and FP32-bulk.  Effective HBM
    bandwidth = 480 B (VAC's dataflow: per sweep the predictor reads w and
    writes w1, the corrector reads w1 and w and writes w; 5*ndim*nw*8 B) per
    cell-update / time."""
    from vac_amd import vacgpu

    method = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.8)
    res = {}
    ref = None
    for prec in ("fp64", "fp32bulk"):
        el, w, kt = timed_run([NX, NY], "mhd", dxs, w0, method, prec, steps,
                              ("tvdlf_fused", "reduce"))
        rate = NX * NY * steps / el
        ksweep = kt["tvdlf_fused"]
        r = {"mcell_updates_per_s": round(rate / 1e6, 2), "ms_per_step": round(el / steps * 1e3, 4),
             "effective_hbm_gbs": round(480 * rate / 1e9, 1),
             "effective_hbm_frac": round(480 * rate / 1e9 / HBM_PEAK_GBS, 4),
             "kernel_avg_ms": kt,
             "sweep_kernel_effective_gbs": round(240 * NX * NY / (ksweep * 1e-3) / 1e9, 1) if ksweep else None}
        if prec == "fp64":
            r["sweep_kernel_traffic_bytes"] = measured_traffic("k_tvdlf_fused")
        if ref is None:
            ref = w
        else:
            r["error_vs_fp64"] = errors(w, ref)
        res[prec] = r
    res["fp32bulk_over_fp64"] = round(res["fp32bulk"]["mcell_updates_per_s"]
                                      / res["fp64"]["mcell_updates_per_s"], 3)
    res["workload"] = (f"2D MHD 2048x2048 periodic random-Fourier, TVDLF+minmod, courantpar 0.8, "
                       f"{steps} steps; B_alg 480 B/cell-update")
    return res


def tvdmu_2048(w0, dxs, steps):
    """2D MHD 2048^2, TVD-MUSCL (typefull='tvdmu') + woodward, FP64.  Same
    dataflow as TVDLF: per sweep the predictor reads w and writes w1, the
    corrector reads w and w1 and writes the result, so B_alg is 480 B per
    cell-update (240 B per sweep; the corrector's share is 144 B)."""
    from vac_amd import vacgpu

    method = vacgpu.MethodSpec(typefull="tvdmu", limiter="woodward", courantpar=0.8)
    el, w, kt = timed_run([NX, NY], "mhd", dxs, w0, method, "fp64", steps,
                          ("tvdlf_pred", "tvdlf_corr", "bound", "reduce"))
    rate = NX * NY * steps / el
    kc = kt["tvdlf_corr"]
    return {"mcell_updates_per_s": round(rate / 1e6, 2), "ms_per_step": round(el / steps * 1e3, 4),
            "effective_hbm_frac": round(480 * rate / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_avg_ms": kt,
            "corrector_effective_gbs": round(144 * NX * NY / (kc * 1e-3) / 1e9, 1) if kc else None,
            "workload": (f"2D MHD 2048x2048 periodic random-Fourier, TVD-MUSCL+woodward, "
                         f"courantpar 0.8, {steps} steps; B_alg 480 B/cell-update")}


def tvdlf_unsplit_2048(w0, dxs, steps):
    """2D MHD 2048^2, TVDLF + minmod with dimsplit=F (one advect over both
    directions per step; the limiter reads the step's start), FP64.  VAC's
    dataflow per step: the predictor reads w and writes w1, the corrector
    reads w1, w and wold and writes w: 6*nw*8 = 288 B per cell-update."""
    from vac_amd import vacgpu

    method = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.8, unsplit=True)
    el, w, kt = timed_run([NX, NY], "mhd", dxs, w0, method, "fp64", steps,
                          ("tvdlf_pred", "tvdlf_corr", "bound", "reduce"))
    rate = NX * NY * steps / el
    return {"mcell_updates_per_s": round(rate / 1e6, 2), "ms_per_step": round(el / steps * 1e3, 4),
            "effective_hbm_frac": round(288 * rate / 1e9 / HBM_PEAK_GBS, 4),
            "kernel_avg_ms": kt,
            "workload": (f"2D MHD 2048x2048 periodic random-Fourier, TVDLF+minmod, dimsplit=F, "
                         f"courantpar 0.8, {steps} steps; B_alg 288 B/cell-update")}


def ot_projection_2048(steps=20):
    """The reference's own Orszag-Tang recipe (Man/par.html:1205-1232: TVD +
    woodward, divbfix=F, divbwave=F, nproc=1, procpar=0,0.01,1, periodic) at
    2048^2, FP64: B is projected (CG Poisson solve) before every step.
    Reported beside the same run without processing; the CG kernels stream
    88 B per cell and iteration (Laplacian 16, update 48, direction 24)."""
    from vac_amd import vacgpu
    from vac_amd.grid import widths_1d
    from vac_amd.initial import orszag_tang

    w0, coords = orszag_tang(NX, GAMMA)
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvd", limiter="woodward", divbfix=False,
                               divbwave=False, courantpar=0.8)
    procpar = (0.0, 0.01, 1.0)
    res = {}
    for nproc1 in (1, 0):
        g = vacgpu.Grid([NX, NX], "mhd", GAMMA, dxs)
        g.put_mesh(w0)
        g.set_process(nproc1, procpar)
        g.run(method, 3)
        g.put_mesh(w0)
        g.set_time(0, 0.0)
        g.set_process(nproc1, procpar)
        t0 = time.perf_counter()
        n, _ = g.run(method, steps)
        el = time.perf_counter() - t0
        if n != steps:
            raise RuntimeError(f"ran {n} of {steps} steps")
        r = {"mcell_updates_per_s": round(NX * NX * steps / el / 1e6, 2),
             "ms_per_step": round(el / steps * 1e3, 4)}
        if nproc1:
            g.profile(True)
            g.run(method, 5)
            r["projection_avg_ms"] = round(g.kernel_time("projectb")[0], 4)
            g.profile(False)
            info, it, resid = g.process(1)
            r["cg_iterations_next_projection"] = it
            r["cg_info"] = info
            r["nerror"] = [int(v) for v in g.scalars()["nerror"]]
        g.close()
        res["with_projection" if nproc1 else "without_processing"] = r
    res["workload"] = (f"Orszag-Tang 2048x2048 periodic, gamma 5/3, TVD+woodward, divbfix=F, "
                       f"divbwave=F, nproc=1, procpar=0,0.01,1, courantpar 0.8, {steps} steps")
    return res


def mhd3d_256(steps=100):
    """3D MHD 256^3, TVDLF + minmod, courantpar 0.6: FP64 and FP32-bulk,
    with the FP32-bulk error per variable against the FP64 result (the FP64
    path is bitwise equal to the oracle at every tested size)."""
    from vac_amd import vacgpu
    from vac_amd.grid import widths_1d
    from vac_amd.initial import random_fourier

    n = 256
    w0, coords = random_fourier((n, n, n), SEED + 3, GAMMA)
    dxs = [widths_1d(c) for c in coords]
    method = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.6)
    res = {}
    ref = None
    for prec in ("fp64", "fp32bulk"):
        el, w, kt = timed_run([n, n, n], "mhd", dxs, w0, method, prec, steps,
                              ("tvdlf_pred", "tvdlf_corr", "bound", "reduce"))
        rate = n ** 3 * steps / el
        r = {"mcell_updates_per_s": round(rate / 1e6, 2), "ms_per_step": round(el / steps * 1e3, 4),
             "effective_hbm_frac": round(960 * rate / 1e9 / HBM_PEAK_GBS, 4),
             "kernel_avg_ms": kt}
        if ref is None:
            ref = w
        else:
            r["error_vs_fp64"] = errors(w, ref)
        res[prec] = r
    res["fp32bulk_over_fp64"] = round(res["fp32bulk"]["mcell_updates_per_s"]
                                      / res["fp64"]["mcell_updates_per_s"], 3)
    res["workload"] = (f"3D MHD 256^3 periodic random-Fourier (B = curl A), TVDLF+minmod, "
                       f"courantpar 0.6, {steps} steps; B_alg 960 B/cell-update")
    return res


def snapshot_overhead(w0, coords, dxs, method, steps=400, every=100):
    """The headline workload with a VAC snapshot every `every` steps: written
    while the device steps (vac_amd.snapshots, asynchronous device snapshots)
    against the same snapshots taken with a synchronous get_state; the file
    goes to /dev/null, so the figures are the step loop's cost of saving."""
    import os as _os

    from vac_amd import snapshots, vacgpu
    from vac_amd.vacfile import write_snapshot

    def fresh():
        g = vacgpu.Grid([NX, NY], "mhd", GAMMA, dxs)
        g.put_mesh(w0)
        return g

    g = fresh()
    g.run(method, 5)  # warm-up: first launches, snapshot buffers
    g.snapshot_start(0)
    g.snapshot_fetch(0)
    g.close()
    g = fresh()
    t0 = time.perf_counter()
    g.run(method, steps)
    plain = time.perf_counter() - t0
    g.close()
    g = fresh()
    t0 = time.perf_counter()
    snapshots.run_with_snapshots(g, method, steps, every, _os.devnull, coords, "mhd", GAMMA,
                                 save_initial=False)
    asyn = time.perf_counter() - t0
    g.close()
    g = fresh()
    t0 = time.perf_counter()
    for _ in range(steps // every):
        g.run(method, every)
        w = g.get_mesh()
        it, t = g.get_time()
        write_snapshot(_os.devnull, snapshots.snapshot_record(w, it, t, coords, "mhd", GAMMA))
    sync = time.perf_counter() - t0
    g.close()
    return {"steps": steps, "every": every, "snapshots": steps // every,
            "ms_per_step_no_snapshots": round(plain / steps * 1e3, 4),
            "ms_per_step_async_snapshots": round(asyn / steps * 1e3, 4),
            "ms_per_step_sync_get_state": round(sync / steps * 1e3, 4),
            "async_overhead": round(asyn / plain - 1.0, 4),
            "sync_overhead": round(sync / plain - 1.0, 4),
            "workload": f"2048^2 TVD headline, a VAC snapshot record every {every} steps (to /dev/null)"}


def batched():
    """512 uneven 2D MHD grids, TVDLF + minmod, 100 steps, FP64 and FP32-bulk."""
    from vac_amd import serve as sv

    insts = [i.prepare() for i in sv.batched_instances(512, seed=4)]
    res = {}
    for prec in ("fp64", "fp32bulk"):
        r = sv.run_batched(insts, steps=100, precision=prec)
        res[prec] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
    res["workload"] = ("512 x 2D MHD, sides log-uniform [64,512] (non-square), gamma in {5/3,1.4}, "
                       "TVDLF+minmod, 100 steps each, one batch, hipGraph replay, autotuned tiles")
    return res


def serving():
    """Open-loop Poisson arrivals at 50, 200 and 1000 requests/s."""
    from vac_amd import serve as sv

    res = {}
    cache = {}
    for rate in (50, 200, 1000, 4000):
        reqs = sv.make_requests(rate, SERVE_REQUESTS, seed=rate, quantum=8)
        sv.prepare_shared(reqs, cache)
        print(f"[bench] serving {rate}/s ({len(reqs)} requests)", file=sys.stderr, flush=True)
        r = sv.serve(reqs, budget=32 << 30)
        res[f"{rate}_per_s"] = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in r.items()}
    res["workload"] = (f"{SERVE_REQUESTS} requests per rate: 2D HD or MHD (50/50), side log-uniform "
                       "[32,1024] rounded to a multiple of 8 (one initial state per side and "
                       "physics), t_end U[0.05,0.5], FP64 or FP32-bulk (50/50), TVDLF+minmod, "
                       "admission by HBM footprint (32 GiB budget), continuous batching; a "
                       "request completes when its final state is in host memory")
    return res


if __name__ == "__main__":
    main()
