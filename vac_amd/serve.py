"""Batched and served solves on one GPU.

``run_batched`` advances many independent 2D TVDLF instances of uneven
sizes together (one vacgpu_batch, each step replayed from a hipGraph, tile
shapes from the autotuner) and reports the aggregate throughput.

``serve`` is a continuous-batching request loop: requests arrive on an
open-loop Poisson clock; each is admitted as soon as its HBM footprint fits
the budget (FIFO), its state then stays resident on the device and is
stepped together with every other admitted request; when a request reaches
its end time (VAC's tmax with tmaxexact) its state is copied back and its
memory reused.  Latency = completion - arrival, per request.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass

import numpy as np

from . import vacgpu
from .grid import widths_1d
from .initial import random_fourier

METHOD = vacgpu.MethodSpec(typefull="tvdlf", limiter="minmod", courantpar=0.8)


def log_uniform_int(rng, lo, hi):
    return int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))


@dataclass
class Instance:
    nx: int
    ny: int
    physics: str
    gamma: float
    seed: int
    precision: str = "fp64"
    t_end: float = 1e99
    arrival: float = 0.0
    w: np.ndarray = None
    dxs: list = None

    def prepare(self):
        self.w, coords = random_fourier((self.nx, self.ny), self.seed, self.gamma,
                                        physics=self.physics)
        self.dxs = [widths_1d(c) for c in coords]
        return self

    @property
    def cells(self):
        return self.nx * self.ny


def batched_instances(count=512, seed=4, lo=64, hi=512):
    """Instances of the batching workload: sides log-uniform in [lo, hi]
    (independently, so non-square), gamma in {5/3, 1.4}, MHD."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx = log_uniform_int(rng, lo, hi)
        ny = log_uniform_int(rng, lo, hi)
        gamma = 5.0 / 3.0 if rng.uniform() < 0.5 else 1.4
        out.append(Instance(nx, ny, "mhd", gamma, 1000 + i))
    return out


def run_batched(instances, steps=100, precision="fp64", budget=64 << 30, graph=True, tune=True):
    """Advance every instance `steps` steps in one batch; returns metrics."""
    b = vacgpu.Batch(budget, len(instances) + 8)
    slots = []
    for inst in instances:
        slots.append(b.add(inst.w, [inst.nx, inst.ny], inst.physics, inst.gamma, inst.dxs, METHOD,
                           precision=precision, itmax=steps))
    if tune:
        b.tune()
    rows, seglen = b.get_tiles()
    b.graph(graph)
    # one untimed step would change the states; instead time from the admitted
    # states, after a synchronisation
    b.sync()
    t0 = time.perf_counter()
    b.step(steps)
    b.sync()
    elapsed = time.perf_counter() - t0
    cells = sum(i.cells for i in instances)
    done = sum(1 for s in slots if b.status(s)["done"] or b.status(s)["it"] == steps)
    its = [b.status(s)["it"] for s in slots]
    b.close()
    return {"instances": len(instances), "steps": steps, "precision": precision,
            "cell_updates": cells * steps, "seconds": elapsed,
            "mcell_updates_per_s": cells * steps / elapsed / 1e6,
            "all_reached_itmax": done == len(instances) and min(its) == steps,
            "tile_rows_per_class": rows, "fused_seglen_per_class": seglen, "graph": graph}


def make_requests(rate, count, seed=7, lo=32, hi=1024, quantum=1):
    """Open-loop Poisson arrivals: 50/50 HD or MHD, square side log-uniform in
    [lo, hi] (rounded to a multiple of quantum), end time uniform in
    [0.05, 0.5], tier FP64 or FP32-bulk."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        n = log_uniform_int(rng, lo, hi)
        if quantum > 1:
            n = max(lo, int(round(n / quantum)) * quantum)
        physics = "mhd" if rng.uniform() < 0.5 else "hd"
        precision = "fp64" if rng.uniform() < 0.5 else "fp32bulk"
        t_end = rng.uniform(0.05, 0.5)
        out.append(Instance(n, n, physics, 5.0 / 3.0, 5000 + i, precision=precision,
                            t_end=t_end, arrival=t))
    return out


def prepare_shared(requests, cache=None):
    """Initial states shared by requests of the same side and physics (the
    state of the first such request's seed): bounds the set-up time of long
    request streams.  Returns the cache for reuse."""
    cache = {} if cache is None else cache
    for r in requests:
        key = (r.nx, r.ny, r.physics)
        if key not in cache:
            cache[key] = r.prepare()
        src = cache[key]
        r.w, r.dxs = src.w, src.dxs
    return cache


def serve(requests, budget=32 << 30, capacity=1024, steps_per_poll=2, keep_states=False,
          graph=True):
    """Serve the requests (already prepared) in arrival order; returns metrics.

    The loop keeps one batch of steps queued on the device while the host
    admits arrivals and reads the previous status snapshot: admission,
    removal, status and result copies are asynchronous (vacgpu_batch), so
    the host blocks only on the snapshot of the steps before the ones still
    queued.  A request completes when its result (the final state) has
    arrived in host memory; latency = completion - arrival.
    """
    b = vacgpu.Batch(budget, capacity)
    b.set_tiles([4, 4, 4, 4])
    b.graph(graph)
    queue = collections.deque()
    active = {}
    fetching = {}
    latencies = []
    results = {}
    cell_updates = 0
    peak_active = 0
    nxt = 0
    inflight = 0
    t_admit = t_wait = t_poll = 0.0
    polls = 0
    start = time.perf_counter()

    def take_snapshot(wait):
        nonlocal inflight, t_wait, t_poll, polls
        tw = time.perf_counter()
        got = b.snapshot_wait(wait)
        tp = time.perf_counter()
        t_wait += tp - tw
        if not got:
            return
        inflight -= 1
        polls += 1
        for slot in b.finished():
            st = b.status(slot)
            r = active.pop(slot)
            fetching[b.fetch_start(slot)] = (r, st)
            b.remove(slot)
        t_poll += time.perf_counter() - tp

    while nxt < len(requests) or queue or active or fetching or inflight:
        now = time.perf_counter() - start
        while nxt < len(requests) and requests[nxt].arrival <= now:
            queue.append(requests[nxt])
            nxt += 1
        ta = time.perf_counter()
        while queue:
            r = queue[0]
            try:
                slot = b.add(r.w, [r.nx, r.ny], r.physics, r.gamma, r.dxs, METHOD,
                             precision=r.precision, tmax=r.t_end, tmaxexact=True)
            except vacgpu.Batch.BudgetFull:
                break
            queue.popleft()
            active[slot] = r
        t_admit += time.perf_counter() - ta
        peak_active = max(peak_active, len(active))
        if active:
            b.step(steps_per_poll)
            b.snapshot()
            inflight += 1
        if inflight >= 2 or (inflight and not active):
            take_snapshot(True)
        elif inflight:
            take_snapshot(False)
        tf = time.perf_counter()
        for tk in list(fetching):
            w = b.fetch_ready(tk, wait=False)
            if w is None:
                continue
            r, st = fetching.pop(tk)
            done_at = time.perf_counter() - start
            cell_updates += r.cells * st["it"]
            latencies.append(done_at - r.arrival)
            if keep_states:
                results[id(r)] = (w, st)
        t_poll += time.perf_counter() - tf
        if not active and not inflight and nxt < len(requests) and not queue:
            if fetching:
                continue
            time.sleep(max(0.0, min(0.001, requests[nxt].arrival - (time.perf_counter() - start))))
    total = time.perf_counter() - start
    b.close()
    lat = np.array(latencies)
    span = requests[-1].arrival - requests[0].arrival if len(requests) > 1 else 0.0
    out = {"requests": len(requests), "seconds": total,
           "arrival_span_s": span,
           "p50_ms": float(np.percentile(lat, 50) * 1e3),
           "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "mean_ms": float(lat.mean() * 1e3),
           "mcell_updates_per_s": cell_updates / total / 1e6,
           "offered_mcell_updates_per_s": cell_updates / span / 1e6 if span > 0 else None,
           "peak_resident": peak_active, "polls": polls,
           "host_s": {"admit": round(t_admit, 4), "wait_for_device": round(t_wait, 4),
                      "status_fetch_remove": round(t_poll, 4)}}
    if keep_states:
        out["results"] = results
    return out
