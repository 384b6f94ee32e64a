"""Serving scenario runs on one GPU: requests arrive (Poisson), are admitted
by device-memory footprint under a budget, join one running batch
(ensemble.Batch) at the next step boundary, keep their state on the device
until they reach their t_end, and leave the batch; freed memory admits the
requests that wait.  Each result equals the same run alone
(tests/test_gpu_batch.py).  Reported: time-to-result percentiles and
cell-stage updates/s, beside a FIFO server of plain runs (one request at a
time through the single-run device driver), whose service times are
measured by running every request alone.

Requests follow the serving workload: arrivals at rate lambda per second
over a window of wall time, a grid side log-uniform in [64, 2048] cells
(dx = 10 m), t_end uniform in a range (physical seconds), the other
parameters drawn as for the batched ensemble."""
from __future__ import annotations

import time
from typing import List

import numpy as np

from . import init, workloads
from .ensemble import Batch
from .solver import Solver

# first admission estimate, device bytes per cell of a one-class run: q, the
# right-hand side and the layout buffer (3 x 5 doubles), maxima (3), deposit
# (1), bed slopes (2), ring fields (3), mask, lists and source bytes (about
# 12 bytes).  Once a run exists the server prices requests at the largest
# bytes per cell a handle has reported (imex_get_footprint) and accounts each
# admitted run at its handle's real footprint.
BYTES_PER_CELL = 8 * (15 + 3 + 1 + 2 + 3) + 12


class Request:
    __slots__ = ("idx", "arrival", "cfg", "static", "q0", "footprint", "admit", "finish",
                 "solver", "cells_stage_updates", "steps", "t_end", "q", "t", "over_budget")

    def __init__(self, idx, arrival, cfg, static, q0):
        self.idx, self.arrival, self.cfg, self.static, self.q0 = idx, arrival, cfg, static, q0
        self.footprint = BYTES_PER_CELL * cfg.nx * cfg.ny
        self.admit = self.finish = None
        self.solver = None
        self.cells_stage_updates = 0
        self.steps = 0
        self.t_end = cfg.t_end
        self.q = None   # final state, when the server keeps results
        self.t = None
        self.over_budget = False


def make_requests(rate: float, duration: float, seed: int = 5, t_end=(60.0, 600.0),
                  side=(64, 2048), n_solid: int = 1) -> List[Request]:
    """Poisson arrivals over `duration` seconds; inputs built on the host
    (the request's payload) before any clock starts."""
    rng = np.random.Generator(np.random.PCG64(5000 + seed))
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        m = 100000 * seed + len(out)
        cfg, B = workloads.config4_member(m, min_side=side[0], max_side=side[1],
                                          t_end=float(rng.uniform(*t_end)), n_solid=n_solid)
        st, q0 = init.setup(cfg, B)
        out.append(Request(len(out), t, cfg, st, q0))
    return out


def serve(reqs: List[Request], budget_bytes: float = 200e9, chunk: int = 20, log=None,
          keep: bool = False):
    """Continuous batching at step boundaries.  Returns the wall time from the
    first arrival to the last result; fills admit/finish/steps per request
    (and the final state and t when keep).

    Admission is in arrival order (no request is overtaken, so none starves).
    A request is admitted while the footprints of the running runs plus its
    own fit the budget; a request larger than the whole budget is admitted
    alone once nothing runs (r.over_budget is set), so the queue always
    drains.  The head request is priced before its handle exists at the
    largest bytes per cell seen so far; once created, the run is accounted at
    its handle's real device footprint."""
    say = log or (lambda *a: None)
    # every request of a stream has the stream's n_solid (a batch takes one)
    b = Batch(n_solid=reqs[0].cfg.n_solid if reqs else 1)
    waiting, running = [], []
    pending = sorted(reqs, key=lambda r: r.arrival)
    used = 0.0
    per_cell = float(BYTES_PER_CELL)
    t0 = time.perf_counter()
    nxt = 0
    while nxt < len(pending) or waiting or running:
        now = time.perf_counter() - t0
        while nxt < len(pending) and pending[nxt].arrival <= now:
            waiting.append(pending[nxt])
            nxt += 1
        # admission by footprint, in arrival order
        while waiting:
            r = waiting[0]
            est = per_cell * r.cfg.nx * r.cfg.ny
            if used + est > budget_bytes:
                if running:
                    break
                r.over_budget = est > budget_bytes
            waiting.pop(0)
            r.solver = Solver(r.cfg, r.static, r.q0)
            r.footprint = r.solver.footprint()
            per_cell = max(per_cell, r.footprint / (r.cfg.nx * r.cfg.ny))
            b.add(r.solver, r.t_end)
            r.admit = time.perf_counter() - t0
            used += r.footprint
            running.append(r)
        if not running:
            if nxt < len(pending):   # idle until the next arrival
                time.sleep(max(0.0, pending[nxt].arrival - (time.perf_counter() - t0)))
            continue
        b.run(chunk)
        done, steps, _ = b.status()
        fin = [r for r, d in zip(list(b.members), done) if d]
        if fin:
            now = time.perf_counter() - t0
            byid = {id(r.solver): r for r in running}
            for s in fin:
                r = byid[id(s)]
                b.remove(s)
                a, n = s.counters()
                r.cells_stage_updates = 2 * a
                r.steps = n
                r.finish = now
                if keep:
                    r.q, r.t = s.q, s.t
                s.close()
                r.solver = None
                used -= r.footprint
                running.remove(r)
    wall = time.perf_counter() - t0
    b.close()
    say(f"served {len(reqs)} requests in {wall:.2f} s")
    return wall


def fifo_service_times(reqs: List[Request], log=None):
    """Each request alone through the single-run device driver, to the step
    count the batch took for it (the same run); returns service seconds.
    Reports its progress every 50 requests through log."""
    say = log or (lambda *a: None)
    out = []
    for i, r in enumerate(reqs):
        s = Solver(r.cfg, r.static, r.q0)
        s.run_steps(0)
        t1 = time.perf_counter()
        s.run_steps(int(r.steps), t_end=r.t_end, record_dt=False)
        out.append(time.perf_counter() - t1)
        s.close()
        if (i + 1) % 50 == 0 or i + 1 == len(reqs):
            say(f"FIFO service times: {i + 1} of {len(reqs)} requests, {sum(out):.1f} s")
    return np.array(out)


def fifo_completion(arrivals, service):
    """One-at-a-time server in arrival order: C_i = max(A_i, C_{i-1}) + S_i."""
    c, out = 0.0, []
    for a, s in zip(arrivals, service):
        c = max(a, c) + s
        out.append(c)
    return np.array(out)


def bench_serving(rate: float, duration: float, t_end=(60.0, 600.0), side=(64, 2048),
                  budget_bytes: float = 200e9, seed: int = 5, log=None, n_solid: int = 1):
    say = log or (lambda *a: None)
    t = time.perf_counter()
    reqs = make_requests(rate, duration, seed=seed, t_end=t_end, side=side, n_solid=n_solid)
    say(f"{len(reqs)} requests generated in {time.perf_counter() - t:.1f} s")
    wall = serve(reqs, budget_bytes=budget_bytes, log=say)
    arr = np.array([r.arrival for r in reqs])
    ttr = np.array([r.finish - r.arrival for r in reqs])
    work = float(sum(r.cells_stage_updates for r in reqs))
    svc = fifo_service_times(reqs, log=say)
    cf = fifo_completion(arr, svc)
    ttr_f = cf - arr
    wall_f = float(cf.max()) if len(reqs) else 0.0
    return {
        "rate_per_s": rate, "window_s": duration, "requests": len(reqs),
        "t_end_range_s": list(t_end), "side_range": list(side),
        "budget_GB": budget_bytes / 1e9,
        "p50_s": float(np.percentile(ttr, 50)), "p99_s": float(np.percentile(ttr, 99)),
        "wall_s": wall, "cell_stage_updates_per_s": work / wall,
        "fifo_p50_s": float(np.percentile(ttr_f, 50)), "fifo_p99_s": float(np.percentile(ttr_f, 99)),
        "fifo_wall_s": wall_f, "fifo_cell_stage_updates_per_s": work / wall_f if wall_f else None,
        "fifo_service_total_s": float(svc.sum()),
        "steps_total": int(sum(r.steps for r in reqs)),
    }
