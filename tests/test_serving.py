"""Continuous-batching server: every result equals the same solve alone.

Requests arrive faster than they finish, so several are resident and
stepped together; each result (FP64 and FP32-bulk tiers) must be bitwise
identical to the same initial state advanced alone to the same end time
(tmax with tmaxexact) through the single-grid path.  The stated tolerance is
bitwise equality for both tiers, since a batched instance runs the same
arithmetic as the single grid.  Each result also matches the oracle run on
its own initial state to its end time: bitwise in FP64, within 1e-5 relative
L1 per variable in FP32-bulk.
"""
import numpy as np
import pytest

from oracle_compare import assert_matches_oracle
from vac_amd import serve as sv
from vac_amd import vacgpu

pytestmark = pytest.mark.gpu


def test_served_results_equal_solo_runs():
    reqs = [r.prepare() for r in sv.make_requests(2000.0, 14, seed=11, lo=24, hi=96)]
    print("case serving 14 requests", flush=True)
    out = sv.serve(reqs, budget=2 << 30, keep_states=True)
    assert out["peak_resident"] >= 2
    for r in reqs:
        wb, st = out["results"][id(r)]
        g = vacgpu.Grid([r.nx, r.ny], r.physics, r.gamma, r.dxs, precision=r.precision)
        g.put_mesh(r.w)
        g.run(sv.METHOD, 1 << 20, tmax=r.t_end, tmaxexact=True)
        wa = g.get_mesh()
        it, t = g.get_time()
        g.close()
        assert it == st["it"] and t == st["t"]
        assert np.array_equal(wa, wb), f"request {r.nx}x{r.ny} {r.physics} {r.precision} differs"
        assert_matches_oracle(wb, st["it"], st["t"], r.w, r.dxs, [r.nx, r.ny], r.physics, r.gamma,
                              r.precision, f"request {r.nx}x{r.ny} {r.physics} {r.precision}",
                              tmax=r.t_end, tmaxexact=True)


def test_config5_scale_serving():
    # the serving workload at its own sizes: sides log-uniform in [32, 1024],
    # HD and MHD, both tiers, arriving at 200/s; every request finishes at
    # its end time with no error counted, and a sample -- the largest of each
    # tier and two more -- against the oracle
    reqs = sv.make_requests(200.0, 24, seed=23)
    for r in reqs:
        r.t_end = min(r.t_end, 0.08)  # bounds the oracle's share of the test
        r.prepare()
    print("case config-5 scale serving: sides", sorted(r.nx for r in reqs), flush=True)
    out = sv.serve(reqs, budget=32 << 30, keep_states=True)
    assert out["requests"] == len(reqs)
    for r in reqs:
        wb, st = out["results"][id(r)]
        assert st["done"] and abs(st["t"] - r.t_end) <= 1e-12, (r.nx, st)
        assert list(st["nerror"]) == [0, 0, 0, 0]
    pick = set()
    for prec in ("fp64", "fp32bulk"):
        tier = [r for r in reqs if r.precision == prec]
        pick.add(id(max(tier, key=lambda r: r.nx)))
    pick |= {id(reqs[3]), id(reqs[11])}
    for r in reqs:
        if id(r) not in pick:
            continue
        wb, st = out["results"][id(r)]
        print(f"  oracle check {r.nx}x{r.ny} {r.physics} {r.precision} t_end={r.t_end:.4f}",
              flush=True)
        assert_matches_oracle(wb, st["it"], st["t"], r.w, r.dxs, [r.nx, r.ny], r.physics, r.gamma,
                              r.precision, f"request {r.nx}x{r.ny} {r.physics} {r.precision}",
                              tmax=r.t_end, tmaxexact=True)
