"""Batches of independent runs on one GPU (ensemble members, served
requests) through the imex_batch_* entry points of include/imex_hip.h: one
launch per kernel and step advances every member, each with its own dt, t
and t_end.  A member's result equals that of the same run alone
(Solver.run_steps), which tests/test_gpu_batch.py checks bit for bit."""
from __future__ import annotations

import ctypes
from typing import Dict, List

import numpy as np

from . import native
from .solver import Solver


class Batch:
    """Owns one imex_batch.  Members are Solver objects (created, static
    fields and initial state uploaded); while in the batch their own step
    calls must not be used."""

    def __init__(self, n_solid: int, blocks_per_member: int = 0, precision: int = 0,
                 flux_threads: int = 256, graphs: bool = True):
        """precision 0: FP64 members, 1: members of the FP32 form (every
        member of a batch has the batch's precision).  flux_threads: the flux
        kernel's block, 256 (a hot half tile in two passes per phase) or 512
        (one pass per phase, or two half tiles per block when they outnumber
        the blocks).  graphs: replay captured HIP graphs of the step rather
        than launching its kernels one by one (the results are the same)."""
        self.L = native.lib()
        self.n_solid = n_solid
        self.precision = precision
        b = ctypes.c_void_p()
        rc = self.L.imex_batch_create_form(ctypes.c_int(n_solid), ctypes.c_int(precision),
                                           ctypes.byref(b))
        self.b = b
        self._check(rc, "imex_batch_create_form")
        self._check(self.L.imex_batch_set_blocks(self.b, ctypes.c_int(blocks_per_member)),
                    "imex_batch_set_blocks")
        self.set_launch(flux_threads, graphs)
        self.members: List[Solver] = []
        self.t_end: Dict[int, float] = {}

    def set_launch(self, flux_threads: int = 256, graphs: bool = True):
        self._check(self.L.imex_batch_set_launch(self.b, ctypes.c_int(flux_threads),
                                                 ctypes.c_int(1 if graphs else 0)),
                    "imex_batch_set_launch")
        self.flux_threads = flux_threads
        self.graphs = graphs

    def _err(self) -> str:
        buf = ctypes.create_string_buffer(512)
        if self.b and self.b.value:
            self.L.imex_batch_last_error(self.b, buf, ctypes.c_size_t(512))
        return buf.value.decode(errors="replace")

    def _check(self, rc: int, what: str):
        if rc < 0:
            raise native.NativeError(f"{what} failed ({rc}): {self._err()}")

    def add(self, s: Solver, t_end: float):
        """The run joins at the next step and stops once its t reaches t_end."""
        s._sync_time_to_device()
        self._check(self.L.imex_batch_add(self.b, s.h, ctypes.c_double(t_end)), "imex_batch_add")
        self.members.append(s)
        self.t_end[id(s)] = t_end

    def remove(self, s: Solver):
        self._check(self.L.imex_batch_remove(self.b, s.h), "imex_batch_remove")
        self.members.remove(s)
        del self.t_end[id(s)]
        s._pull_time()

    def run(self, nsteps: int):
        """nsteps steps of every member still short of its t_end (asynchronous)."""
        self._check(self.L.imex_batch_run(self.b, ctypes.c_int(nsteps)), "imex_batch_run")

    def status(self):
        """(done[bool], steps[int], t[float]) per member, in self.members
        order; synchronises."""
        n = len(self.members)
        done = np.zeros(max(n, 1), np.int32)
        steps = np.zeros(max(n, 1), np.int32)
        t = np.zeros(max(n, 1))
        self._check(self.L.imex_batch_status(self.b, ctypes.c_int(n), native._ip(done),
                                             native._ip(steps), native._dp(t)),
                    "imex_batch_status")
        return done[:n].astype(bool), steps[:n], t[:n]

    def run_to_end(self, chunk: int = 50, max_steps: int = 10 ** 7) -> int:
        """Steps until every member is done; returns the batch steps taken."""
        steps = 0
        while steps < max_steps:
            done, _, _ = self.status()
            if done.all():
                break
            self.run(chunk)
            steps += chunk
        return steps

    def close(self):
        if getattr(self, "b", None) is not None and self.b.value:
            self.L.imex_batch_destroy(self.b)
            self.b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_alone(s: Solver, t_end: float, max_steps: int = 10 ** 7) -> int:
    """The same run through the single-run device driver, one step per call
    so that it stops exactly where a batch member stops (t >= t_end)."""
    n = 0
    while s.t < t_end and n < max_steps:
        s.run_steps(1, t_end=t_end, record_dt=False)
        n += 1
    return n


def autotune(batch: Batch, blocks=(0, 4, 8, 16, 32, 64), flux_threads=(256, 512),
             steps: int = 8):
    """Picks the launch shape of a batch on the batch itself: the flux
    kernel's work shape (a hot half tile on 256 threads in two passes per
    phase, or on 512 threads in one pass) and the blocks per member (with
    the member count, the waves per CU of each launch; 0 = the library's
    choice).  Each candidate times `steps` batch steps of the members, which
    advance meanwhile (useful work, nothing is undone), and the fastest is
    kept.  Returns ((blocks, flux_threads), {"blocks/threads": seconds per
    step})."""
    import time
    table = {}
    for ft in flux_threads:
        batch.set_launch(ft, batch.graphs)
        for gx in blocks:
            batch._check(batch.L.imex_batch_set_blocks(batch.b, ctypes.c_int(gx)),
                         "imex_batch_set_blocks")
            batch.status()
            batch.run(1)           # the capture of a new launch shape, outside the clock
            batch.status()
            t0 = time.perf_counter()
            batch.run(steps)
            batch.status()
            table[(gx, ft)] = (time.perf_counter() - t0) / steps
    best = min(table, key=table.get)
    batch.set_launch(best[1], batch.graphs)
    batch._check(batch.L.imex_batch_set_blocks(batch.b, ctypes.c_int(best[0])),
                 "imex_batch_set_blocks")
    return best, {f"{k[0]}/{k[1]}": v for k, v in table.items()}


def autotune_blocks(batch: Batch, candidates=(0, 4, 8, 16, 32, 64), steps: int = 8):
    """autotune over the blocks per member only (the flux shape kept)."""
    best, table = autotune(batch, blocks=candidates, flux_threads=(batch.flux_threads,),
                           steps=steps)
    return best[0], table


def bench_ensemble(n_members: int = 256, t_end: float = 30.0, chunk: int = 25,
                   min_side: int = 128, max_side: int = 1024, log=None, precision: int = 0):
    """The batched ensemble against the same runs one after the other through
    the single-run device driver.  Set-up (DEMs, static fields, uploads) is
    outside both clocks.  Returns a dict of the measurements."""
    import time
    from . import init, workloads
    say = log or (lambda *a: None)
    t_set = time.perf_counter()
    specs = []
    for m in range(n_members):
        cfg, B = workloads.config4_member(m, min_side=min_side, max_side=max_side, t_end=t_end)
        cfg.precision = precision
        st, q0 = init.setup(cfg, B)
        specs.append((cfg, st, q0))
    say(f"host set-up of {n_members} members: {time.perf_counter() - t_set:.1f} s")
    cells = np.array([c.nx * c.ny for c, _, _ in specs])

    # batched
    sol = [Solver(*sp) for sp in specs]
    b = Batch(n_solid=1, precision=precision)
    for s, (cfg, _, _) in zip(sol, specs):
        b.add(s, cfg.t_end)
    b.status()
    t0 = time.perf_counter()
    best, table = autotune(b)
    steps = b.run_to_end(chunk=chunk)
    done, nsteps, tfin = b.status()
    wall_batch = time.perf_counter() - t0
    assert done.all()
    say(f"batched: {wall_batch:.3f} s, blocks per member {best}, max member steps {nsteps.max()}")
    for s in list(sol):
        b.remove(s)
    b.close()
    q_batch = {i: sol[i].q for i in (0, n_members // 2, n_members - 1)}
    for s in sol:
        s.close()

    # sequential: the same member runs, each to its step count, one by one
    wall_seq = 0.0
    same = True
    for i, sp in enumerate(specs):
        s = Solver(*sp)
        s.run_steps(0)
        t1 = time.perf_counter()
        s.run_steps(int(nsteps[i]), t_end=sp[0].t_end, record_dt=False)
        wall_seq += time.perf_counter() - t1
        if i in q_batch:
            same = same and np.array_equal(s.q.view(np.int64), q_batch[i].view(np.int64)) \
                and s.t == tfin[i]
        s.close()
    say(f"sequential: {wall_seq:.3f} s")
    return {
        "members": n_members, "t_end_s": t_end, "cells_total": int(cells.sum()),
        "cells_min": int(cells.min()), "cells_max": int(cells.max()),
        "member_steps_total": int(nsteps.sum()), "member_steps_max": int(nsteps.max()),
        "batch_wall_s": wall_batch, "sequential_wall_s": wall_seq,
        "runs_per_hour_batched": n_members / wall_batch * 3600.0,
        "runs_per_hour_sequential": n_members / wall_seq * 3600.0,
        "speedup": wall_seq / wall_batch,
        "precision": "fp32" if precision else "fp64",
        "blocks_per_member": best[0], "flux_threads": best[1], "hip_graphs": True,
        "autotune_s_per_step": table,
        "batch_equals_alone_checked_members": sorted(q_batch), "batch_equals_alone": bool(same),
        "_member_steps": nsteps.tolist(),
    }
