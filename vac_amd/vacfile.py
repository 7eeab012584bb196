"""VAC binary snapshot files (.ini / .out), read and written from Python.

The layout is the one VAC's own reader and writer use
(src/vacio.t:682-755 ``readfileini_bin``, src/vacio.t:1008-1047
``savefileout_bin``): Fortran unformatted sequential records, each framed by a
4-byte little-endian length before and after the payload.

  record 1  file header, character*79 (``^LENNAME``, src/vacpp.pl:154)
  record 2  it (int32), t (float64), ndim (int32), neqpar (int32), nw (int32)
  record 3  nx(1:ndim) (int32)
  record 4  eqpar(1:neqpar) (float64)
  record 5  variable names, character*79
  record 6  x(ix^S, 1:ndim) (float64, Fortran order)
  record 7.. one record per variable: w(ix^S, iw) (float64, Fortran order)

A file may hold several snapshots back to back; :func:`read_snapshots`
returns all of them, :func:`read_snapshot` the last one (VAC's default for
``snapshotini=0``).

Arrays are returned C-ordered with the spatial axes reversed, so that the
x index is the fastest one in memory: ``x[..., idim]`` has shape
``(nz, ny, nx)`` and ``w[iw]`` has shape ``(nz, ny, nx)`` (2D: ``(ny, nx)``).
This is the same memory order as VAC's column-major ``w(ix1, ix2, iw)``.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field

import numpy as np

LENNAME = 79


@dataclass
class Snapshot:
    header: str
    it: int
    t: float
    ndim: int
    nw: int
    nx: tuple
    eqpar: np.ndarray
    varnames: str
    x: np.ndarray          # (..., ndim), spatial axes reversed (x fastest)
    w: np.ndarray          # (nw, ...), spatial axes reversed (x fastest)
    gencoord: bool = field(default=False)


def _rec(buf: bytes) -> bytes:
    n = len(buf)
    return struct.pack("<i", n) + buf + struct.pack("<i", n)


def _fstr(s: str) -> bytes:
    b = s.encode("ascii")
    if len(b) > LENNAME:
        raise ValueError(f"string longer than {LENNAME} characters: {s!r}")
    return b.ljust(LENNAME, b" ")


def _records(snap: Snapshot):
    """The records of one snapshot as buffers, the large ones without copies."""
    ndim = snap.ndim
    nx = tuple(int(n) for n in snap.nx)
    if len(nx) != ndim:
        raise ValueError("len(nx) != ndim")
    shape = tuple(reversed(nx))
    x = np.ascontiguousarray(snap.x, dtype="<f8")
    w = np.ascontiguousarray(snap.w, dtype="<f8")
    if x.shape != shape + (ndim,):
        raise ValueError(f"x shape {x.shape} != {shape + (ndim,)}")
    if w.shape != (snap.nw,) + shape:
        raise ValueError(f"w shape {w.shape} != {(snap.nw,) + shape}")
    ndimout = -ndim if snap.gencoord else ndim
    small = [
        _fstr(snap.header),
        struct.pack("<idiii", int(snap.it), float(snap.t), ndimout, len(snap.eqpar),
                    int(snap.nw)),
        struct.pack("<" + "i" * ndim, *nx),
        np.asarray(snap.eqpar, dtype="<f8").tobytes(),
        _fstr(snap.varnames),
    ]
    for b in small:
        yield _rec(b)
    # x(ix^S, idim) for idim=1..ndim, column-major == planes of x fastest,
    # then w(ix^S, iw) one record per variable
    big = [np.ascontiguousarray(np.moveaxis(x, -1, 0))] + [w[iw] for iw in range(snap.nw)]
    for a in big:
        n = struct.pack("<i", a.nbytes)
        yield n
        yield memoryview(a).cast("B")
        yield n


def snapshot_bytes(snap: Snapshot) -> bytes:
    """Serialise one snapshot in VAC's binary layout."""
    return b"".join(bytes(b) for b in _records(snap))


def write_snapshot(path: str, snap: Snapshot, append: bool = False) -> None:
    """Write (or append) one snapshot; the arrays go to the file unbuffered
    and uncopied."""
    with open(path, "ab" if append else "wb") as f:
        for b in _records(snap):
            f.write(b)


class _Reader:
    def __init__(self, data: bytes):
        self.data = data
        self.pos = 0

    def eof(self) -> bool:
        return self.pos >= len(self.data)

    def record(self) -> bytes:
        if self.pos + 4 > len(self.data):
            raise EOFError
        (n,) = struct.unpack_from("<i", self.data, self.pos)
        start = self.pos + 4
        end = start + n
        if end + 4 > len(self.data):
            raise ValueError("truncated Fortran record")
        (n2,) = struct.unpack_from("<i", self.data, end)
        if n2 != n:
            raise ValueError("corrupt Fortran record markers")
        self.pos = end + 4
        return self.data[start:end]


def read_snapshots(path: str) -> list:
    with open(path, "rb") as f:
        data = f.read()
    r = _Reader(data)
    snaps = []
    while not r.eof():
        header = r.record().decode("ascii").rstrip()
        it, t, ndimout, neqpar, nw = struct.unpack("<idiii", r.record())
        ndim = abs(ndimout)
        nx = struct.unpack("<" + "i" * ndim, r.record())
        eqpar = np.frombuffer(r.record(), dtype="<f8").copy()
        if len(eqpar) != neqpar:
            raise ValueError("eqpar record length mismatch")
        varnames = r.record().decode("ascii").rstrip()
        shape = tuple(reversed(nx))
        xr = np.frombuffer(r.record(), dtype="<f8").reshape((ndim,) + shape)
        x = np.ascontiguousarray(np.moveaxis(xr, 0, -1))
        w = np.empty((nw,) + shape)
        for iw in range(nw):
            w[iw] = np.frombuffer(r.record(), dtype="<f8").reshape(shape)
        snaps.append(Snapshot(header, it, t, ndim, nw, tuple(nx), eqpar,
                              varnames, x, w, gencoord=ndimout < 0))
    return snaps


def read_snapshot(path: str, index: int = -1) -> Snapshot:
    snaps = read_snapshots(path)
    if not snaps:
        raise ValueError(f"no snapshot in {path}")
    return snaps[index]
