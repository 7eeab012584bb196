"""Asynchronous device snapshots and VAC snapshot files (save and restart).

A snapshot started on the device after step k must hold exactly the state
(ghost cells included), it and t of step k, however many steps are enqueued
after it before it is fetched.  A run that appends snapshots to a VAC .out
file while it steps must write the same records a synchronous run would, and
VAC itself must be able to restart from such a file: the reference
executable continuing from the GPU's snapshot and a GPU grid restarted from
the same file must agree bit for bit.
"""
import os

import numpy as np
import pytest

from vac_amd import snapshots, vacgpu, vacfile
from vac_amd.grid import widths_1d
from vac_amd.initial import random_fourier

G = 5 / 3
METHOD = vacgpu.MethodSpec(typefull="tvd", limiter="woodward", courantpar=0.8)


def test_snapshot_record_matches_vac_names_and_round_trips(tmp_path):
    # the names record is VAC's (as the reference's own files carry it), and a
    # record written here reads back unchanged
    from run_ref import hd_varnames, mhd_varnames
    assert snapshots.varnames("mhd", 2) == mhd_varnames(2)[1]
    assert snapshots.varnames("mhd", 3) == mhd_varnames(3)[1]
    assert snapshots.varnames("hd", 2) == hd_varnames(2)[1]
    w, coords = random_fourier((12, 9), 3, G)
    rec = snapshots.snapshot_record(w, 17, 0.25, coords, "mhd", G)
    p = str(tmp_path / "a.out")
    vacfile.write_snapshot(p, rec)
    vacfile.write_snapshot(p, snapshots.snapshot_record(2 * w, 18, 0.5, coords, "mhd", G),
                           append=True)
    s = vacfile.read_snapshots(p)
    assert [x.it for x in s] == [17, 18] and [x.t for x in s] == [0.25, 0.5]
    assert np.array_equal(s[0].w, w) and np.array_equal(s[1].w, 2 * w)
    assert s[0].varnames.strip() == mhd_varnames(2)[1]


def _grid(w0, coords, physics="mhd"):
    g = vacgpu.Grid([len(c) for c in coords], physics, G, [widths_1d(c) for c in coords])
    g.put_mesh(w0)
    return g


@pytest.mark.gpu
def test_async_snapshot_holds_the_state_it_was_taken_at():
    w0, coords = random_fourier((70, 45), 11, G)
    print("case async snapshot", flush=True)
    a = _grid(w0, coords)
    a.run(METHOD, 5)
    a.snapshot_start(0)
    a.run(METHOD, 7)
    a.snapshot_start(1)
    a.run(METHOD, 3)
    s1 = a.snapshot_fetch(1, mesh=False)
    s0 = a.snapshot_fetch(0, mesh=False)
    a.close()
    b = _grid(w0, coords)
    b.run(METHOD, 5)
    r0 = (b.get_state(),) + b.get_time()
    b.run(METHOD, 7)
    r1 = (b.get_state(),) + b.get_time()
    b.close()
    for s, r in ((s0, r0), (s1, r1)):
        assert np.array_equal(s[0], r[0])
        assert (s[1], s[2]) == (r[1], r[2])


@pytest.mark.gpu
def test_run_with_snapshots_writes_the_synchronous_records(tmp_path):
    w0, coords = random_fourier((64, 40), 12, G)
    path = str(tmp_path / "run.out")
    print("case run_with_snapshots", flush=True)
    a = _grid(w0, coords)
    written = snapshots.run_with_snapshots(a, METHOD, 12, 5, path, coords, "mhd", G)
    a.close()
    snaps = vacfile.read_snapshots(path)
    assert [s.it for s in snaps] == [0, 5, 10, 12] == [it for it, _ in written]
    b = _grid(w0, coords)
    ref = [(b.get_mesh(),) + b.get_time()]
    for n in (5, 5, 2):
        b.run(METHOD, n)
        ref.append((b.get_mesh(),) + b.get_time())
    b.close()
    for s, (w, it, t) in zip(snaps, ref):
        assert (s.it, s.t) == (it, t)
        assert np.array_equal(s.w, w)


@pytest.mark.gpu
def test_vac_restarts_from_a_gpu_snapshot(tmp_path):
    from run_ref import have_ref, problem_par, run_ref
    if not have_ref("vac_mhd22_132"):
        pytest.skip("reference executable not built (oracle/build_ref.sh)")
    w0, coords = random_fourier((40, 30), 13, G)
    path = str(tmp_path / "gpu.out")
    print("case vac restart from gpu snapshot", flush=True)
    a = _grid(w0, coords)
    snapshots.run_with_snapshots(a, METHOD, 6, 3, path, coords, "mhd", G)
    a.close()
    last = vacfile.read_snapshot(path)
    assert last.it == 6
    par = problem_par(6, "rho m1 m2 e b1 b2", 6 + 8, typefull="tvd", limiter="woodward",
                      every_step=True)
    r = run_ref("vac_mhd22_132", vacfile.snapshot_bytes(last), par)
    g, snap = snapshots.restart(lambda: vacgpu.Grid([40, 30], "mhd", G,
                                                    [widths_1d(c) for c in coords]), path)
    n, t = g.run(METHOD, 8, record_t=True)
    w = g.get_mesh()
    g.close()
    # VAC saves after every step of the restarted run (its it = 0 output does
    # not recur at it = 6): t after steps 7..14
    assert [s.it for s in r.snapshots] == list(range(7, 15))
    assert np.array_equal(np.array([s.t for s in r.snapshots]), t[1:])
    assert np.array_equal(r.snapshots[-1].w, w)
