"""Device-resident runs with VAC snapshot files written while the device steps.

VAC's main loop (src/vac.t:55-106) calls ``savefile`` when an output is due;
``savefileout_bin`` (src/vacio.t:1008-1047) writes the host copy of ``w`` as
one binary record set (see :mod:`vac_amd.vacfile`).  With the state resident
on the device that would mean a synchronous download every time.  Here a
snapshot is taken on the device in stream order (``vacgpu_snapshot_start``:
a device copy, then an asynchronous copy to pinned host memory), the next
steps start at once, and a host thread fetches and writes the snapshot while
they run.  Two slots alternate, so one file write overlaps one chunk of
steps.

A file written here is a regular VAC ``.out`` file: VAC (or
:func:`restart`) can read any snapshot of it as an initial state
(``filenameini``, ``snapshotini``, src/vacio.t:682-755).
"""
from __future__ import annotations

import numpy as np

from .initial import coords_array
from .vacfile import Snapshot, read_snapshot, write_snapshot


def varnames(physics: str, ndim: int) -> str:
    """The names record of VAC's files: coordinates, variables, equation
    parameters (as VAC's vacpar modules define them)."""
    xs = " ".join("xyz"[:ndim])
    ms = " ".join(f"m{d + 1}" for d in range(ndim))
    if physics == "mhd":
        bs = " ".join(f"b{d + 1}" for d in range(ndim))
        return f"{xs} rho {ms} e {bs} gamma eta"
    return f"{xs} rho {ms} e gamma"


def snapshot_record(w_mesh: np.ndarray, it: int, t: float, coords: list, physics: str,
                    gamma: float, header: str = "vacgpu", x: np.ndarray | None = None) -> Snapshot:
    """One snapshot record of a mesh state (w[iw] with x fastest); x, the
    full coordinate array, may be passed in to be built only once."""
    ndim = len(coords)
    eqpar = np.array([gamma, 0.0]) if physics == "mhd" else np.array([gamma])
    return Snapshot(f"{header}_{physics}{ndim}{ndim}", int(it), float(t), ndim, w_mesh.shape[0],
                    tuple(len(c) for c in coords), eqpar, varnames(physics, ndim),
                    coords_array(coords) if x is None else x, np.ascontiguousarray(w_mesh))


def run_with_snapshots(grid, method, nsteps: int, every: int, path: str, coords: list,
                       physics: str, gamma: float, header: str = "vacgpu",
                       save_initial: bool = True) -> list:
    """Advance `grid` nsteps with the device-resident loop, appending a
    snapshot to the VAC file `path` every `every` steps (and of the initial
    state when save_initial).  Returns the (it, t) of every snapshot written,
    in file order.

    Each snapshot is started on the device in stream order right after its
    chunk of steps; a worker thread fetches it and appends it to the file
    while the main thread runs the next chunk (the HIP calls release the
    interpreter lock).  A slot is reused only after its previous write."""
    from concurrent.futures import ThreadPoolExecutor

    written = []
    jobs = [None, None]
    x = coords_array(coords)

    def write(slot, first):
        w, it, t = grid.snapshot_fetch(slot)
        write_snapshot(path, snapshot_record(w, it, t, coords, physics, gamma, header, x),
                       append=not first)
        written.append((it, t))

    with ThreadPoolExecutor(max_workers=1) as pool:
        slot = 0
        first = True

        def take():
            nonlocal slot, first
            if jobs[slot] is not None:
                jobs[slot].result()
            grid.snapshot_start(slot)
            jobs[slot] = pool.submit(write, slot, first)  # one worker: file order kept
            first = False
            slot = 1 - slot

        if save_initial:
            take()
        done = 0
        while done < nsteps:
            n = min(every, nsteps - done)
            grid.run(method, n)
            take()
            done += n
        for j in jobs:
            if j is not None:
                j.result()
    return written


def restart(grid_factory, path: str, index: int = -1):
    """A grid holding snapshot `index` of a VAC file, with its it and t set,
    ready to continue as VAC does after reading an initial file."""
    snap = read_snapshot(path, index)
    grid = grid_factory()
    grid.put_mesh(snap.w)
    grid.set_time(snap.it, snap.t)
    return grid, snap
