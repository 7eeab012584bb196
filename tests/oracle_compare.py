"""Comparison of a GPU result with the oracle run on the same initial state
(shared by the batch and serving tests).

The bar: FP64 results are bitwise the oracle's, with the same it and t;
FP32-bulk results are within 1e-5 relative L1 per conserved variable of the
FP64 oracle at the same it (and their t within the same relative bar).
"""
import numpy as np

import vacref

TOL_FP32 = 1e-5


def oracle(w, dxs, nx, physics, gamma, steps=None, tmax=1e99, tmaxexact=False, bc="periodic",
           typefull="tvdlf", limiter="minmod", courantpar=0.8):
    p = vacref.make_params(len(nx), physics, list(nx), gamma, typefull=typefull, limiter=limiter,
                           courantpar=courantpar, bc=bc, tmax=tmax, tmaxexact=tmaxexact)
    wr, t, st = vacref.run(p, w, dxs, steps if steps is not None else 1 << 20)
    return wr, len(t) - 1, t[-1], st


def assert_matches_oracle(wb, it_b, t_b, w0, dxs, nx, physics, gamma, prec, what, tmax=1e99,
                          tmaxexact=False, bc="periodic", **method):
    wr, it_r, t_r, _ = oracle(w0, dxs, nx, physics, gamma,
                              steps=None if tmax < 1e99 else it_b, tmax=tmax,
                              tmaxexact=tmaxexact, bc=bc, **method)
    if prec == "fp64":
        assert (it_b, t_b) == (it_r, t_r), f"{what}: it/t {it_b} {t_b} vs oracle {it_r} {t_r}"
        assert np.array_equal(wb, wr), f"{what}: differs from the oracle"
        return
    assert abs(t_b - t_r) <= TOL_FP32 * t_r, f"{what}: t {t_b} vs oracle {t_r}"
    if tmax >= 1e99:
        assert it_b == it_r
    for iw in range(wr.shape[0]):
        e = np.abs(wb[iw] - wr[iw]).sum() / np.abs(wr[iw]).sum()
        assert e <= TOL_FP32, f"{what}: variable {iw} rel L1 {e:.3e} against the oracle"
