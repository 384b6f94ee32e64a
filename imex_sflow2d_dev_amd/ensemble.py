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

    def __init__(self, n_solid: int, blocks_per_member: int = 64):
        self.L = native.lib()
        self.n_solid = n_solid
        b = ctypes.c_void_p()
        rc = self.L.imex_batch_create(ctypes.c_int(n_solid), ctypes.byref(b))
        self.b = b
        self._check(rc, "imex_batch_create")
        self._check(self.L.imex_batch_set_blocks(self.b, ctypes.c_int(blocks_per_member)),
                    "imex_batch_set_blocks")
        self.members: List[Solver] = []
        self.t_end: Dict[int, float] = {}

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
        """(done[bool], t[float]) per member, in self.members order; synchronises."""
        n = len(self.members)
        done = np.zeros(max(n, 1), np.int32)
        t = np.zeros(max(n, 1))
        self._check(self.L.imex_batch_status(self.b, ctypes.c_int(n), native._ip(done),
                                             native._dp(t)), "imex_batch_status")
        return done[:n].astype(bool), t[:n]

    def run_to_end(self, chunk: int = 50, max_steps: int = 10 ** 7) -> int:
        """Steps until every member is done; returns the batch steps taken."""
        steps = 0
        while steps < max_steps:
            done, _ = self.status()
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
