"""The quotient and the square root of the device kernels against the
compiler's IEEE division and sqrt.

Every division by a positive quantity on the hot path (densities, cell
widths, squared sound speeds, the Roe-mean sqrt(rho)) goes through one
routine: a Newton-Raphson refined reciprocal, shared by all quotients with
the same denominator, and one Markstein residual correction.  It must give
the same double as a correctly rounded division over the operand range a
physical state produces, so FP64 results stay bitwise equal to VAC's.  The
operands cover wide exponent ranges, significands with all bits set,
quotients at and next to exact values, and zero numerators of both signs.

Every square root takes the compiler's own sequence without its scaling and
special-value steps wherever those do nothing, and sqrt() elsewhere: it must
give the same double for every input, special values included.
"""
import ctypes as C

import numpy as np
import pytest

from vac_amd import vacgpu

pytestmark = pytest.mark.gpu


def operands(n, seed):
    rng = np.random.default_rng(seed)
    b = np.exp2(rng.uniform(-300.0, 300.0, n)) * rng.uniform(1.0, 2.0, n)
    a = np.exp2(rng.uniform(-600.0, 600.0, n)) * rng.uniform(1.0, 2.0, n)
    a = np.where(rng.random(n) < 0.5, -a, a)
    # significands with every bit set, and powers of two
    k = n // 8
    ones = np.nextafter(np.exp2(rng.integers(-200, 200, k).astype(float) + 1.0), 0.0)
    b[:k] = ones
    a[k:2 * k] = np.nextafter(np.exp2(rng.integers(-200, 200, k).astype(float) + 1.0), 0.0)
    b[2 * k:3 * k] = np.exp2(rng.integers(-200, 200, k).astype(float))
    # exact and nearly exact quotients: a = b * small integer, +- 1 ulp
    q = rng.integers(1, 1 << 20, k).astype(float)
    a[3 * k:4 * k] = b[3 * k:4 * k] * q
    a[4 * k:5 * k] = np.nextafter(b[4 * k:5 * k] * q, np.inf)
    a[5 * k:6 * k] = np.nextafter(b[5 * k:6 * k] * q, -np.inf)
    # the magnitudes of the sweep kernels: O(1e-3..1e3) over O(1e-2..1e2)
    b[6 * k:7 * k] = rng.uniform(1e-2, 1e2, k)
    a[6 * k:7 * k] = rng.uniform(-1e3, 1e3, k)
    a[7 * k:7 * k + 64] = 0.0
    a[7 * k + 64:7 * k + 128] = -0.0
    return np.ascontiguousarray(a), np.ascontiguousarray(b)


def test_quotient_matches_ieee_division():
    lib = vacgpu.load_library()
    n = 1 << 22
    a, b = operands(n, 20261203)
    qk = np.empty(n)
    qi = np.empty(n)
    p = lambda x: x.ctypes.data_as(C.POINTER(C.c_double))
    print("case quotient self-test", n, flush=True)
    assert lib.vacgpu_check_quotients(p(a), p(b), n, p(qk), p(qi)) == 0
    # the compiler's division is correctly rounded: numpy's quotient
    assert np.array_equal(qi.view(np.uint64), (a / b).view(np.uint64))
    # every operand here lies in the stated range
    tiny = np.finfo(np.float64).tiny
    q = np.abs(a / b)
    assert np.all((b >= tiny) & (1.0 / b >= tiny) & ((a == 0) | (np.abs(a) >= 2.0**-969))
                  & ((a == 0) | (q >= tiny)) & np.isfinite(q))
    bad = np.nonzero(qk.view(np.uint64) != qi.view(np.uint64))[0]
    assert bad.size == 0, [(a[i], b[i], qk[i], qi[i]) for i in bad[:5]]


def root_operands(n, seed):
    rng = np.random.default_rng(seed)
    x = np.exp2(rng.uniform(-1000.0, 1000.0, n)) * rng.uniform(1.0, 2.0, n)
    k = n // 8
    # significands with every bit set, and powers of two
    x[:k] = np.nextafter(np.exp2(rng.integers(-900, 900, k).astype(float) + 1.0), 0.0)
    x[k:2 * k] = np.exp2(rng.integers(-1000, 1000, k).astype(float))
    # perfect squares and their neighbours (roots at and next to exact values)
    r = rng.uniform(1.0, 2.0**26, k).round() * np.exp2(rng.integers(-300, 300, k).astype(float))
    x[2 * k:3 * k] = r * r
    x[3 * k:4 * k] = np.nextafter(r * r, np.inf)
    x[4 * k:5 * k] = np.nextafter(r * r, -np.inf)
    # the magnitudes of the sweep kernels
    x[5 * k:6 * k] = rng.uniform(1e-6, 1e3, k)
    # around the scaling threshold 2^-767, subnormals and the special values
    t = 2.0**-767
    m = 6 * k
    for v in (t, np.nextafter(t, 0.0), np.nextafter(t, 1.0), 0.0, -0.0, np.inf, -np.inf, np.nan,
              -1.0, -t, 5e-324, 2.2e-308, np.finfo(np.float64).max, np.finfo(np.float64).tiny):
        x[m] = v
        m += 1
    x[m:m + 4096] = np.exp2(rng.uniform(-1074.0, -760.0, 4096))
    return np.ascontiguousarray(x)


def test_root_matches_ieee_sqrt():
    lib = vacgpu.load_library()
    n = 1 << 22
    x = root_operands(n, 20261204)
    rk = np.empty(n)
    ri = np.empty(n)
    p = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))
    print("case square-root self-test", n, flush=True)
    assert lib.vacgpu_check_roots(p(x), n, p(rk), p(ri)) == 0
    # the compiler's square root is correctly rounded: numpy's root (NaN for x < 0)
    with np.errstate(invalid="ignore"):
        ref = np.sqrt(x)
    fin = ~np.isnan(ref)
    assert np.array_equal(ri[fin].view(np.uint64), ref[fin].view(np.uint64))
    assert np.all(np.isnan(ri[~fin]))
    same = (rk.view(np.uint64) == ri.view(np.uint64)) | (np.isnan(rk) & np.isnan(ri))
    bad = np.nonzero(~same)[0]
    assert bad.size == 0, [(x[i], rk[i], ri[i]) for i in bad[:5]]
