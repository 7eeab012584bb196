"""Batched and continuously batched grids (vacgpu_batch) on the GPU.

An instance advanced in a batch must be bitwise identical to the same grid
advanced alone through the single-grid path -- whatever the mix of sizes,
physics, gamma and arithmetic tier around it, whether the step runs eagerly
or from a replayed hipGraph, and whenever it was admitted -- and match the
oracle run on its own initial state: bitwise in FP64, within 1e-5 relative
L1 per variable in FP32-bulk.
"""
import numpy as np
import pytest

from oracle_compare import assert_matches_oracle as _cmp
from vac_amd import vacgpu
from vac_amd.grid import cell_centres, widths_1d
from vac_amd.initial import random_fourier

pytestmark = pytest.mark.gpu

METHOD = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.8)

MIX = [
    # (nx, ny, physics, gamma, precision, seed)
    (64, 48, "mhd", 5 / 3, "fp64", 1),
    (40, 72, "hd", 1.4, "fp64", 2),
    (96, 96, "mhd", 1.4, "fp64", 3),
    (128, 64, "hd", 5 / 3, "fp64", 4),
    (33, 65, "mhd", 5 / 3, "fp32bulk", 5),
    (200, 150, "mhd", 5 / 3, "fp64", 6),
    (72, 40, "hd", 1.4, "fp32bulk", 7),
    (260, 300, "hd", 1.4, "fp64", 8),
]


def state(nx, ny, physics, gamma, seed):
    w, coords = random_fourier((nx, ny), seed, gamma, physics=physics)
    return w, [widths_1d(c) for c in coords]


def alone(w, dxs, nx, ny, physics, gamma, precision, steps=None, tmax=1e99, tmaxexact=False):
    g = vacgpu.Grid([nx, ny], physics, gamma, dxs, precision=precision)
    g.put_mesh(w)
    n, t = g.run(METHOD, steps if steps is not None else 1 << 20, tmax=tmax, tmaxexact=tmaxexact,
                 record_t=False)
    out = g.get_mesh()
    it, tt = g.get_time()
    g.close()
    return out, it, tt


def assert_matches_oracle(wb, it_b, t_b, w0, dxs, nx, ny, physics, gamma, prec, what, **kw):
    _cmp(wb, it_b, t_b, w0, dxs, [nx, ny], physics, gamma, prec, what, **kw)


def run_batch(batch, entries, max_steps):
    for _ in range(max_steps):
        batch.step(1)
    batch.sync()
    return {s: (batch.get_mesh(s), batch.status(s)) for s in entries}


def test_batch_matches_each_grid_alone():
    b = vacgpu.Batch(8 << 30, 64)
    entries = {}
    for (nx, ny, ph, g, prec, seed) in MIX:
        w, dxs = state(nx, ny, ph, g, seed)
        s = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=12)
        entries[s] = (w, dxs, nx, ny, ph, g, prec)
    print("case batch mix of", len(entries), flush=True)
    res = run_batch(b, entries, 14)
    for s, (w, dxs, nx, ny, ph, g, prec) in entries.items():
        wb, st = res[s]
        assert st["done"] and st["it"] == 12
        wa, it, t = alone(w, dxs, nx, ny, ph, g, prec, steps=12)
        assert it == 12 and t == st["t"], (s, t, st["t"])
        assert np.array_equal(wa, wb), f"instance {s} ({nx}x{ny} {ph} {prec}) differs"
        assert_matches_oracle(wb, st["it"], st["t"], w, dxs, nx, ny, ph, g, prec,
                              f"instance {s} ({nx}x{ny} {ph} {prec})")
    b.close()


def test_graph_replays_match_eager():
    out = []
    for use_graph in (False, True):
        b = vacgpu.Batch(4 << 30, 16)
        slots = []
        for (nx, ny, ph, g, prec, seed) in MIX[:5]:
            w, dxs = state(nx, ny, ph, g, seed)
            slots.append(b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=9))
        b.graph(use_graph)
        print(f"case graph={use_graph}", flush=True)
        snaps = []
        for rep in range(4):  # several replays, with synchronisation between them
            b.step(2)
            b.sync()
            snaps.append([b.get_mesh(s) for s in slots])
        out.append(snaps)
        b.close()
    for rep in range(4):
        for a, c in zip(out[0][rep], out[1][rep]):
            assert np.array_equal(a, c), f"replay {rep} differs from the eager step"


def test_continuous_admission_matches_alone():
    b = vacgpu.Batch(4 << 30, 16)
    pending = list(MIX[:6])
    live = {}
    steps = 0
    print("case continuous admission", flush=True)
    while pending or live:
        # admit one instance every third step (so admissions land on both parities)
        if pending and steps % 3 == 0:
            nx, ny, ph, g, prec, seed = pending.pop(0)
            w, dxs = state(nx, ny, ph, g, seed)
            itmax = 6 + seed
            s = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=itmax)
            live[s] = (w, dxs, nx, ny, ph, g, prec, itmax)
        b.step(1)
        steps += 1
        b.sync()
        for s in list(live):
            st = b.status(s)
            if st["done"]:
                w, dxs, nx, ny, ph, g, prec, itmax = live.pop(s)
                wb = b.get_mesh(s)
                b.remove(s)
                wa, it, t = alone(w, dxs, nx, ny, ph, g, prec, steps=itmax)
                assert st["it"] == itmax and st["t"] == t
                assert np.array_equal(wa, wb), f"instance admitted late differs ({nx}x{ny})"
                assert_matches_oracle(wb, st["it"], st["t"], w, dxs, nx, ny, ph, g, prec,
                                      f"instance admitted late ({nx}x{ny} {ph} {prec})")
        assert steps < 200
    b.close()


def test_tmax_exact_per_instance():
    b = vacgpu.Batch(2 << 30, 8)
    cases = []
    for (nx, ny, ph, g, prec, seed), tmax in zip(MIX[:3], (0.05, 0.08, 0.031)):
        w, dxs = state(nx, ny, ph, g, seed)
        s = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, tmax=tmax, tmaxexact=True)
        cases.append((s, w, dxs, nx, ny, ph, g, prec, tmax))
    print("case tmaxexact", flush=True)
    for _ in range(200):
        b.step(1)
    b.sync()
    for (s, w, dxs, nx, ny, ph, g, prec, tmax) in cases:
        st = b.status(s)
        assert st["done"]
        assert abs(st["t"] - tmax) <= 1e-12
        wa, it, t = alone(w, dxs, nx, ny, ph, g, prec, tmax=tmax, tmaxexact=True)
        assert it == st["it"] and t == st["t"]
        assert np.array_equal(wa, b.get_mesh(s))
        assert_matches_oracle(b.get_mesh(s), st["it"], st["t"], w, dxs, nx, ny, ph, g, prec,
                              f"tmaxexact instance ({nx}x{ny})", tmax=tmax, tmaxexact=True)
    b.close()


def test_budget_admission():
    fp = vacgpu.Batch.footprint([128, 128], "mhd")
    b = vacgpu.Batch(int(fp * 2.5), 8)
    w, dxs = state(128, 128, "mhd", 5 / 3, 3)
    b.add(w, [128, 128], "mhd", 5 / 3, dxs, METHOD)
    b.add(w, [128, 128], "mhd", 5 / 3, dxs, METHOD)
    with pytest.raises(vacgpu.Batch.BudgetFull):
        b.add(w, [128, 128], "mhd", 5 / 3, dxs, METHOD)
    assert b.bytes_used() == 2 * fp
    b.close()


def test_autotune_keeps_results():
    b = vacgpu.Batch(4 << 30, 16)
    entries = {}
    for (nx, ny, ph, g, prec, seed) in MIX:
        w, dxs = state(nx, ny, ph, g, seed)
        s = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=5)
        entries[s] = (w, dxs, nx, ny, ph, g, prec)
    rows = b.tune()
    print("case autotune rows", rows, flush=True)
    assert all(r in (1, 2, 4, 8) for r in rows)
    res = run_batch(b, entries, 6)
    for s, (w, dxs, nx, ny, ph, g, prec) in entries.items():
        wa, it, t = alone(w, dxs, nx, ny, ph, g, prec, steps=5)
        assert np.array_equal(wa, res[s][0])
    b.close()


def test_fused_and_tile_sweeps_agree():
    # eligible instances run the fused predictor-corrector sweep; forcing the
    # predictor + corrector tiles for every instance must give the same bits,
    # and a 'cont' instance and a mixed-limiter one ride along in both
    method_mixed = vacgpu.MethodSpec(typefull="tvdlf", limiter=["minmod", "woodward"] * 4,
                                     courantpar=0.8)
    res = []
    precs = []
    for fused in (True, False):
        b = vacgpu.Batch(4 << 30, 16)
        b.set_kernels(fused, [8, 16, 5, 32])
        slots = []
        precs = []
        for j, (nx, ny, ph, g, prec, seed) in enumerate(MIX[:6]):
            w, dxs = state(nx, ny, ph, g, seed)
            bc = "cont" if j == 2 else "periodic"
            m = method_mixed if j == 3 else METHOD
            slots.append(b.add(w, [nx, ny], ph, g, dxs, m, bc=bc, precision=prec, itmax=7))
            precs.append(prec)
        print(f"case batch fused={fused}", flush=True)
        for _ in range(8):
            b.step(1)
        b.sync()
        res.append([(b.get_mesh(s), b.status(s)["t"]) for s in slots])
        b.close()
    # the FP64 instances of the plain method against the oracle
    for j, (nx, ny, ph, g, prec, seed) in enumerate(MIX[:6]):
        if j == 3 or prec != "fp64":
            continue
        w, dxs = state(nx, ny, ph, g, seed)
        wa, ta = res[0][j]
        assert_matches_oracle(wa, 7, ta, w, dxs, nx, ny, ph, g, prec, f"fused instance {j}",
                              bc="cont" if j == 2 else "periodic")
    for (wa, ta), (wb, tb), prec in zip(res[0], res[1], precs):
        if prec == "fp64":
            assert ta == tb
            assert np.array_equal(wa, wb)
        else:  # the fused FP32-bulk sweep has its own increment form: the tier's bar
            assert abs(ta - tb) <= 1e-5 * tb
            for iw in range(wa.shape[0]):
                assert np.abs(wa[iw] - wb[iw]).sum() <= 1e-5 * np.abs(wb[iw]).sum()


def test_tune_reports_segment_lengths():
    b = vacgpu.Batch(4 << 30, 16)
    for (nx, ny, ph, g, prec, seed) in MIX:
        w, dxs = state(nx, ny, ph, g, seed)
        b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=3)
    b.tune()
    rows, sl = b.get_tiles()
    print("case tune tiles", rows, sl, flush=True)
    assert all(r in (1, 2, 4, 8) for r in rows)
    assert all(x in (8, 16, 32, 64) for x in sl)
    b.close()


def test_config4_scale_batch():
    # 512 instances, sides log-uniform in [64, 512] (non-square), gamma in
    # {5/3, 1.4}, TVDLF + minmod, both tiers: every instance reaches itmax
    # with no error counted and sum(rho) conserved to its tier's rounding
    # (periodic faces; rho's update is in flux form); a sample -- the
    # largest, the smallest and random ones of both tiers -- against the
    # oracle
    from vac_amd import serve as sv

    steps = 20
    insts = [i.prepare() for i in sv.batched_instances(512, seed=4)]
    rng = np.random.default_rng(44)
    b = vacgpu.Batch(64 << 30, 512)
    slots = []
    for k, inst in enumerate(insts):
        prec = "fp64" if k % 2 == 0 else "fp32bulk"
        slots.append((b.add(inst.w, [inst.nx, inst.ny], "mhd", inst.gamma, inst.dxs, METHOD,
                            precision=prec, itmax=steps), prec))
    print(f"case config-4 scale: {len(slots)} instances, {steps} steps", flush=True)
    b.graph(True)
    b.step(steps + 1)
    b.sync()
    for (s, prec), inst in zip(slots, insts):
        st = b.status(s)
        assert st["done"] and st["it"] == steps, (inst.nx, inst.ny, st)
        assert list(st["nerror"]) == [0, 0, 0, 0], (inst.nx, inst.ny, st["nerror"])
        wb = b.get_mesh(s)
        # FP64: the flux differences telescope up to FP64 rounding; FP32-bulk
        # rounds each cell's flux difference to FP32 before the FP64 sum
        r0, r1 = inst.w[0].sum(), wb[0].sum()
        tol = 1e-12 if prec == "fp64" else 1e-10
        assert abs(r1 - r0) <= tol * r0, (inst.nx, inst.ny, prec, r0, r1)
    cells = [i.nx * i.ny for i in insts]
    pick = {int(np.argmax(cells)), int(np.argmin(cells)), int(np.argmax(cells[1::2])) * 2 + 1}
    pick |= set(int(k) for k in rng.choice(len(insts), 5, replace=False))
    for k in sorted(pick):
        inst, (s, prec) = insts[k], slots[k]
        st = b.status(s)
        print(f"  oracle check instance {k}: {inst.nx}x{inst.ny} {prec}", flush=True)
        assert_matches_oracle(b.get_mesh(s), st["it"], st["t"], inst.w, inst.dxs, inst.nx,
                              inst.ny, "mhd", inst.gamma, prec, f"instance {k}")
    b.close()


def test_ixg_and_mesh_admission_agree():
    # vacgpu_batch_add takes the ixG state (ghost values from the caller,
    # refilled on the device), vacgpu_batch_add_mesh the mesh only
    import ctypes as C
    nx, ny, ph, g, prec, seed = MIX[0]
    w, dxs = state(nx, ny, ph, g, seed)
    b = vacgpu.Batch(2 << 30, 8)
    s_mesh = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=7)
    cfg = vacgpu.make_config([nx, ny], ph, g, precision=prec)
    m = METHOD.to_c(ph)
    wg = np.ascontiguousarray(np.pad(w, [(0, 0), (2, 2), (2, 2)], constant_values=np.nan))
    dx0 = np.ascontiguousarray(dxs[0])
    dx1 = np.ascontiguousarray(dxs[1])
    slot = C.c_int()
    dp = C.POINTER(C.c_double)
    rc = b.lib.vacgpu_batch_add(b.h, C.byref(cfg), C.byref(m), dx0.ctypes.data_as(dp),
                                dx1.ctypes.data_as(dp), wg.ctypes.data_as(dp), None, 0, 0.0, 7,
                                1e99, 0, C.byref(slot))
    assert rc == 0
    b.meta[slot.value] = ((nx, ny), w.shape[0])
    print("case ixG and mesh admission", flush=True)
    res = run_batch(b, [s_mesh, slot.value], 8)
    assert np.array_equal(res[s_mesh][0], res[slot.value][0])
    wa, it, t = alone(w, dxs, nx, ny, ph, g, prec, steps=7)
    assert np.array_equal(wa, res[s_mesh][0])
    b.close()


def test_pipelined_slot_reuse_matches_alone():
    # two slots, six instances: each is admitted into a slot freed by a
    # removal while later steps are still queued, fetched asynchronously and
    # compared with the same grid alone
    b = vacgpu.Batch(2 << 30, 2)
    b.graph(True)
    pending = [MIX[i] + (5 + i,) for i in (0, 1, 2, 4, 6, 3)]
    active, fetching, done = {}, {}, []
    inflight = 0
    print("case pipelined slot reuse", flush=True)
    for _ in range(400):
        while pending and len(active) < 2:
            nx, ny, ph, g, prec, seed, itmax = pending[0]
            w, dxs = state(nx, ny, ph, g, seed)
            try:
                s = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=itmax)
            except vacgpu.Batch.BudgetFull:
                break
            pending.pop(0)
            active[s] = (w, dxs, nx, ny, ph, g, prec, itmax)
        if active:
            b.step(2)
            b.snapshot()
            inflight += 1
        if inflight >= 2 or (inflight and not active):
            assert b.snapshot_wait(True)
            inflight -= 1
            for s in b.finished():
                st = b.status(s)
                fetching[b.fetch_start(s)] = (active.pop(s), st)
                b.remove(s)
        for tk in list(fetching):
            wb = b.fetch_ready(tk, wait=False)
            if wb is not None:
                done.append((fetching.pop(tk), wb))
        if not pending and not active and not fetching and not inflight:
            break
    assert len(done) == 6
    for (inst, st), wb in done:
        w, dxs, nx, ny, ph, g, prec, itmax = inst
        assert st["done"] and st["it"] == itmax
        wa, it, t = alone(w, dxs, nx, ny, ph, g, prec, steps=itmax)
        assert t == st["t"]
        assert np.array_equal(wa, wb), f"reused-slot instance {nx}x{ny} {ph} {prec} differs"
    b.close()


TVDMU_MIX = [
    # (nx, ny, physics, gamma, precision, seed, limiter)
    (64, 48, "mhd", 5 / 3, "fp64", 21, "woodward"),
    (40, 72, "hd", 1.4, "fp64", 22, "minmod"),
    (96, 80, "mhd", 1.4, "fp64", 23, "superbee"),
    (130, 66, "mhd", 5 / 3, "fp32bulk", 24, "woodward"),
    (72, 40, "hd", 1.4, "fp32bulk", 25, "minmod"),
]


def test_tvdmu_instances_match_alone():
    # TVD-MUSCL instances (predictor tiles, corrector strips and columns,
    # their two-wave Courant fold) beside TVDLF instances in one batch: each
    # bitwise the same grid alone, and against the oracle
    b = vacgpu.Batch(4 << 30, 32)
    b.graph(True)
    entries = {}
    for (nx, ny, ph, g, prec, seed, lim) in TVDMU_MIX:
        w, dxs = state(nx, ny, ph, g, seed)
        m = vacgpu.MethodSpec(typefull="tvdmu", limiter=lim, courantpar=0.8)
        s = b.add(w, [nx, ny], ph, g, dxs, m, precision=prec, itmax=9 + seed % 3)
        entries[s] = (w, dxs, nx, ny, ph, g, prec, m, 9 + seed % 3, lim)
    for (nx, ny, ph, g, prec, seed) in MIX[:2]:
        w, dxs = state(nx, ny, ph, g, seed)
        s = b.add(w, [nx, ny], ph, g, dxs, METHOD, precision=prec, itmax=10)
        entries[s] = (w, dxs, nx, ny, ph, g, prec, METHOD, 10, "minmod")
    print("case TVD-MUSCL and TVDLF instances in one batch", flush=True)
    for _ in range(13):
        b.step(1)
    b.sync()
    for s, (w, dxs, nx, ny, ph, g, prec, m, itmax, lim) in entries.items():
        st = b.status(s)
        wb = b.get_mesh(s)
        assert st["done"] and st["it"] == itmax, (s, st)
        ga = vacgpu.Grid([nx, ny], ph, g, dxs, precision=prec)
        ga.put_mesh(w)
        ga.run(m, itmax, record_t=False)
        wa = ga.get_mesh()
        it, t = ga.get_time()
        ga.close()
        what = f"{m.typefull} {lim} {nx}x{ny} {ph} {prec}"
        assert it == itmax and t == st["t"], what
        assert np.array_equal(wa, wb), f"{what} differs from the grid alone"
        assert_matches_oracle(wb, st["it"], st["t"], w, dxs, nx, ny, ph, g, prec, what,
                              typefull=m.typefull, limiter=lim)
    b.close()


def test_tvdmu_admission_rules():
    b = vacgpu.Batch(1 << 30, 4)
    nx, ny, ph, g, prec, seed = MIX[0]
    w, dxs = state(nx, ny, ph, g, seed)
    with pytest.raises(vacgpu.VacGpuError, match="entropy"):
        b.add(w, [nx, ny], ph, g, dxs, vacgpu.MethodSpec(typefull="tvdmu", limiter="woodward",
                                                        typeentropy="harten"))
    with pytest.raises(vacgpu.VacGpuError, match="'tvdlf' and 'tvdmu'"):
        b.add(w, [nx, ny], ph, g, dxs, vacgpu.MethodSpec(typefull="tvd", limiter="woodward"))
    b.close()
