"""Per-cell phase clocks of the update kernel on the 2048 x 2048
radial-source run (two solid classes, settling and entrainment, FP64; or
another named workload, --workload).

Runs the phase-clock build of the library (_lib/libimexhip_phases.so, made by
`python -m imex_sflow2d_dev_amd.build --phase-clocks`: the same kernels with
shader-clock readings at the phase boundaries of each cell's update, see
PhaseSlot in csrc/imex_physics.h) through the device-side driver, records
one step at each of the given step numbers and prints, per step, the
distribution over the step's work cells of the cycles from the cell's start
to each phase boundary and of the work counts, and the phase split of the
slowest cell.  A lane pair's cycles are the pair's (both lanes run
together); a cell's cycles include the waiting for the other pairs of its
wavefront, which run in step.

    python -m imex_sflow2d_dev_amd.phase_clocks [--steps 600,1500,2500,3500,4200]
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os

import numpy as np

from . import build, native

SLOTS = ["start", "stage2", "newton0", "newton1", "stage2_end", "sed0", "sed1", "end",
         "newton_its", "jacobian_cycles", "line_search_cycles", "line_search_rounds",
         "settling_cycles", "settling_its", "powers_cycles", "complex_residuals"]


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--steps", default="600,1500,2500,3500,4200")
    ap.add_argument("--workload", default="c2", help="c2 (default) or c1")
    ap.add_argument("--json", default="", help="also write the summaries to this file")
    args = ap.parse_args()
    steps = sorted(int(x) for x in args.steps.split(","))
    if not os.path.exists(build.LIB_PHASES):
        raise SystemExit(f"{build.LIB_PHASES} is not built "
                         "(python -m imex_sflow2d_dev_amd.build --phase-clocks)")
    native.LIB_PATH = build.LIB_PHASES      # before the first library call
    from . import init, workloads
    from .solver import Solver
    L = ctypes.CDLL(build.LIB_PHASES)
    cfg, B = workloads.by_name(args.workload)
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0, raise_on_error=False)
    done, out = 0, {}
    for target in steps:
        if target - 1 > done:
            s.run_steps(target - 1 - done)
            done = target - 1
        w0, _ = s.work_counters()
        assert L.imex_phase_clocks_enable(ctypes.c_int(1)) == 0
        s.run_steps(1)
        done += 1
        assert L.imex_phase_clocks_enable(ctypes.c_int(0)) == 0
        w1, _ = s.work_counters()
        n = min(w1 - w0, 65536)
        buf = np.zeros((n, 16), np.uint64)
        assert L.imex_phase_clocks_read(buf.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n)) == 0
        b = buf.astype(np.int64)
        seg = {
            "load": b[:, 1],                                    # start -> stage 2
            "q_fv + semi-implicit": b[:, 2] - b[:, 1],
            "Newton solve": b[:, 3] - b[:, 2],
            "assembly checks": b[:, 4] - b[:, 3],
            "sedimentation loads": b[:, 5] - b[:, 4],
            "sedimentation": b[:, 6] - b[:, 5],
            "epilogue + stores": b[:, 7] - b[:, 6],
            "whole cell": b[:, 7],
        }
        cnt = {k: b[:, i] for i, k in enumerate(SLOTS) if 8 <= i <= 15}
        row = {"step": target, "work_cells": int(w1 - w0)}
        row["cycles"] = {k: {"p50": float(np.median(v)), "p99": float(np.percentile(v, 99)),
                             "max": int(v.max())} for k, v in seg.items()}
        row["counts"] = {k: {"p50": float(np.median(v)), "max": int(v.max()),
                             "hist": np.bincount(np.minimum(v, 40)).tolist()
                             if "its" in k or "rounds" in k or "complex" in k else None}
                         for k, v in cnt.items()}
        j = int(np.argmax(b[:, 7]))
        row["slowest_cell"] = {k: int(v[j]) for k, v in list(seg.items()) + list(cnt.items())}
        out[str(target)] = row
        print(json.dumps(row), flush=True)
    s.close()
    if args.json:
        with open(args.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
