"""Initial states for the MHD and HD advect path, in VAC's conserved variables.

Two families are provided, both on the periodic box [0, 2*pi]^ndim with the
uniform cell centres of :func:`vac_amd.grid.cell_centres`:

* :func:`orszag_tang` -- the 2D Orszag-Tang vortex as VACINI's ``wave``
  recipe defines it (Man/inipar.html:302-312): rho = gamma^2, p = gamma,
  v = (-sin y, sin x), B = (-sin y, sin 2x), with the exact gamma^2 and gamma
  instead of the recipe's rounded 2.778 and 1.667.

* :func:`random_fourier` -- a seeded sum of random Fourier modes:
  rho = 1 + 0.2 * sum_8 (1/8) cos(k.x + phi), p = 1,
  each velocity component 0.1 * sum_8 (1/8) cos(k.x + phi) over fresh modes,
  and B derived from a random vector potential (2D: A_z; 3D: A_x, A_y, A_z,
  each 0.1 * sum_8 (1/8) cos(k.x + phi) / |k|) by the same centred difference
  VAC's ``gradient`` uses (src/vacgrid.t:816-823), so the discrete divergence
  VAC's Powell source computes vanishes up to rounding.  Wave vectors are
  drawn uniformly from the integer lattice points with 1 <= |k| <= 8 and the
  phases uniformly from [0, 2*pi), both from a SplitMix64 stream seeded per
  instance.

Conserved variables follow VAC's ``conserve`` (src/vacphys0.t.mhd:19-23):
e = p/(gamma-1) + 0.5*(rho*|v|^2 + |B|^2), m = rho*v.
Variable order is VAC's: MHD (rho, m1..m_ndir, e, b1..b_ndir), HD
(rho, m1..m_ndim, e).  Arrays are (nw, [nz,] ny, nx), x fastest.
"""
from __future__ import annotations

import math

import numpy as np

from .grid import cell_centres

MASK64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 generator (Steele, Lea, Flood 2014)."""

    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, n: int) -> int:
        return min(int(self.uniform() * n), n - 1)


def _lattice(ndim: int, kmax: int = 8) -> list:
    pts = []
    rng = range(-kmax, kmax + 1)
    if ndim == 2:
        for ky in rng:
            for kx in rng:
                if 1 <= kx * kx + ky * ky <= kmax * kmax:
                    pts.append((kx, ky))
    elif ndim == 3:
        for kz in rng:
            for ky in rng:
                for kx in rng:
                    if 1 <= kx * kx + ky * ky + kz * kz <= kmax * kmax:
                        pts.append((kx, ky, kz))
    else:
        raise ValueError("ndim must be 2 or 3")
    return pts


def _mode_sum(rng: SplitMix64, coords: list, lattice: list, amp: float,
              nmodes: int = 8, over_k: bool = False) -> np.ndarray:
    """amp * sum over nmodes of (1/nmodes) cos(k.x + phi) [/|k|].

    cos(k.x + phi) is split into a product of 1D factors: in 2D
    cos(kx x + ky y + phi) = cos(kx x) cos(ky y + phi) - sin(kx x) sin(ky y + phi),
    so the whole sum is one small matrix product (ny x 2M) @ (2M x nx); in
    3D the z and y factors are combined first.
    """
    ndim = len(coords)
    modes = []
    for _ in range(nmodes):
        k = lattice[rng.below(len(lattice))]
        phi = 2.0 * math.pi * rng.uniform()
        wgt = 1.0 / nmodes
        if over_k:
            wgt = wgt / math.sqrt(sum(kk * kk for kk in k))
        modes.append((k, phi, wgt))
    x = coords[0]
    if ndim == 2:
        y = coords[1]
        A = np.empty((len(y), 2 * nmodes))
        B = np.empty((2 * nmodes, len(x)))
        for m, (k, phi, wgt) in enumerate(modes):
            A[:, 2 * m] = wgt * np.cos(k[1] * y + phi)
            A[:, 2 * m + 1] = -wgt * np.sin(k[1] * y + phi)
            B[2 * m] = np.cos(k[0] * x)
            B[2 * m + 1] = np.sin(k[0] * x)
        return amp * (A @ B)
    y, z = coords[1], coords[2]
    ny, nz = len(y), len(z)
    A = np.empty((nz, ny, 2 * nmodes))
    B = np.empty((2 * nmodes, len(x)))
    for m, (k, phi, wgt) in enumerate(modes):
        ph = k[2] * z[:, None] + k[1] * y[None, :] + phi
        A[:, :, 2 * m] = wgt * np.cos(ph)
        A[:, :, 2 * m + 1] = -wgt * np.sin(ph)
        B[2 * m] = np.cos(k[0] * x)
        B[2 * m + 1] = np.sin(k[0] * x)
    return amp * (A.reshape(nz * ny, -1) @ B).reshape(nz, ny, len(x))


def _centred_diff(a: np.ndarray, d: int, dx_d: np.ndarray, ndim: int) -> np.ndarray:
    """0.5*(a(i+1) - a(i-1))/dx(i) along dimension d with periodic wrap."""
    axis = ndim - 1 - d
    ap = np.roll(a, -1, axis=axis)
    am = np.roll(a, 1, axis=axis)
    bshape = [1] * ndim
    bshape[axis] = a.shape[axis]
    return 0.5 * (ap - am) / dx_d.reshape(bshape)


def _conserve(rho, v, b, p, gamma):
    vsq = sum(vc * vc for vc in v)
    e = p / (gamma - 1.0) + 0.5 * (rho * vsq + (sum(bc * bc for bc in b) if b else 0.0))
    m = [rho * vc for vc in v]
    return m, e


def orszag_tang(n: int, gamma: float = 5.0 / 3.0, ny: int | None = None):
    """2D MHD Orszag-Tang state; returns (w[6, ny, nx], coords[x, y])."""
    ny = n if ny is None else ny
    x = cell_centres(n)
    y = cell_centres(ny)
    X = np.broadcast_to(x[None, :], (ny, n))
    Y = np.broadcast_to(y[:, None], (ny, n))
    rho = np.full((ny, n), gamma * gamma)
    p = np.full((ny, n), gamma)
    v = [-np.sin(Y), np.sin(X)]
    b = [-np.sin(Y), np.sin(2.0 * X)]
    m, e = _conserve(rho, v, b, p, gamma)
    w = np.stack([rho, m[0], m[1], e, b[0], b[1]])
    return np.ascontiguousarray(w), [x, y]


def random_fourier(shape_nx: tuple, seed: int, gamma: float = 5.0 / 3.0,
                   physics: str = "mhd"):
    """Seeded random-Fourier state on [0,2pi]^ndim.

    ``shape_nx`` lists the cell counts in dimension order (nx, ny[, nz]).
    Returns (w[nw, ...], coords) with arrays laid out x fastest.
    """
    from .grid import widths_1d

    ndim = len(shape_nx)
    coords = [cell_centres(int(nd)) for nd in shape_nx]
    lattice = _lattice(ndim)
    rng = SplitMix64(seed)
    rho = 1.0 + _mode_sum(rng, coords, lattice, 0.2)
    p = np.ones_like(rho)
    v = [_mode_sum(rng, coords, lattice, 0.1) for _ in range(ndim)]
    if physics == "hd":
        m, e = _conserve(rho, v, [], p, gamma)
        w = np.stack([rho] + m + [e])
        return np.ascontiguousarray(w), coords
    if physics != "mhd":
        raise ValueError("physics must be 'mhd' or 'hd'")
    # widths exactly as VAC computes them, used for the discrete curl
    dxs = [widths_1d(c)[2:-2] for c in coords]
    if ndim == 2:
        az = _mode_sum(rng, coords, lattice, 0.1, over_k=True)
        b = [_centred_diff(az, 1, dxs[1], 2), -_centred_diff(az, 0, dxs[0], 2)]
    else:
        a = [_mode_sum(rng, coords, lattice, 0.1, over_k=True) for _ in range(3)]
        # B = curl A with centred differences
        b = [_centred_diff(a[2], 1, dxs[1], 3) - _centred_diff(a[1], 2, dxs[2], 3),
             _centred_diff(a[0], 2, dxs[2], 3) - _centred_diff(a[2], 0, dxs[0], 3),
             _centred_diff(a[1], 0, dxs[0], 3) - _centred_diff(a[0], 1, dxs[1], 3)]
    m, e = _conserve(rho, v, b, p, gamma)
    w = np.stack([rho] + m + [e] + b)
    return np.ascontiguousarray(w), coords


def coords_array(coords: list) -> np.ndarray:
    """Full coordinate array x[..., idim] (x fastest) for a tensor grid."""
    ndim = len(coords)
    shape = tuple(len(c) for c in reversed(coords))
    out = np.empty(shape + (ndim,))
    for d in range(ndim):
        bshape = [1] * ndim
        bshape[ndim - 1 - d] = len(coords[d])
        out[..., d] = np.broadcast_to(coords[d].reshape(bshape), shape)
    return out
