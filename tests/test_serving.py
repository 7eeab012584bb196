"""Continuous-batching server: every result equals the same solve alone.

Requests arrive faster than they finish, so several are resident and
stepped together; each result (FP64 and FP32-bulk tiers) must be bitwise
identical to the same initial state advanced alone to the same end time
(tmax with tmaxexact) through the single-grid path.  The stated tolerance is
bitwise equality for both tiers, since a batched instance runs the same
arithmetic as the single grid.
"""
import numpy as np
import pytest

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
