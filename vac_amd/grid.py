"""Cartesian grid geometry as VAC derives it from the cell-centre coordinates.

VAC reads the coordinates of the mesh cells ``ixM`` from the input file and
fills the two ghost layers per side (plus one extra layer used only for the
differences) by linear extrapolation from the edge cell and its inner
neighbour, then sets the per-cell width to the centred difference
(src/vacgrid.t:457-530, ``gridsetup1``):

    qx(ixe +- k) = (1 + k) * x(ixe) - k * x(ixf)     k = 1, 2, 3
    dx(i)        = 0.5 * (qx(i+1) - qx(i-1))          i in ixG

On a slab grid whose coordinate ``d`` depends on index ``d`` only, the width
along ``d`` is a 1D array over the ``n_d + 4`` cells of ``ixG``; that is the
form the GPU library takes. The cell volume is the product of the widths in
dimension order, ``(dx1*dx2)*dx3`` (src/vacgrid.t:574-575).

The arithmetic is done in float64 exactly in the order above, so the widths
are bitwise the ones VAC computes from the same coordinates.
"""
from __future__ import annotations

import numpy as np

NGHOST = 2


def widths_1d(xc: np.ndarray) -> np.ndarray:
    """Cell widths over ixG (length n+4) from the n cell-centre coordinates."""
    xc = np.asarray(xc, dtype=np.float64)
    n = xc.shape[0]
    if n < 2:
        raise ValueError("need at least two cells per dimension")
    # extended index e = i + 3 for mesh index i in [-3, n+2]
    q = np.empty(n + 6)
    q[3:3 + n] = xc
    for k in (1, 2, 3):
        # upper layers: ixe = n-1, ixf = n-2
        q[3 + n - 1 + k] = float(1 + k) * xc[n - 1] - float(k) * xc[n - 2]
        # lower layers: ixe = 0, ixf = 1
        q[3 - k] = float(1 + k) * xc[0] - float(k) * xc[1]
    dx = np.empty(n + 4)
    for i in range(-2, n + 2):
        dx[i + 2] = 0.5 * (q[i + 3 + 1] - q[i + 3 - 1])
    return dx


def cell_centres(n: int, length: float = 2.0 * np.pi, origin: float = 0.0) -> np.ndarray:
    """Uniform cell centres x_i = origin + (i + 1/2) * length / n, i = 0..n-1.

    This is the ``domain`` construction of VACINI (src/vacini.t:258-265)
    evaluated in float64 with an exact float64 2*pi.
    """
    h = length / n
    return origin + (np.arange(n, dtype=np.float64) + 0.5) * h


def widths_from_snapshot_x(x: np.ndarray) -> list:
    """Per-dimension widths from a snapshot coordinate array ``x[..., idim]``.

    Raises ValueError if coordinate ``d`` varies along another index (a
    non-tensor grid, which only VAC's generalized-coordinate module handles).
    """
    ndim = x.shape[-1]
    out = []
    for d in range(ndim):
        comp = x[..., d]
        axis = ndim - 1 - d           # arrays are stored with axes reversed
        line = np.moveaxis(comp, axis, -1).reshape(-1, comp.shape[axis])
        if not np.all(line == line[0]):
            raise ValueError(f"coordinate {d + 1} is not a tensor-product grid")
        out.append(widths_1d(line[0]))
    return out
