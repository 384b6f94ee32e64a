"""Throughput of the FP64 IMEX-RK step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2]

Workload (default "c2"): the production-size radial-source flow, 2048 x 2048
cells of 5 m on a seeded fractal cone, two solid classes, sedimentation and
air entrainment on, FP64, a run of 5000 driver steps.  One step is one
driver step of the reference's time loop (src/IMEX_SfloW2D.f90:395-505):
check_solve, timestep, imex_RK_solver (two RK stages),
update_erosion_deposition_cell and the epilogue, all run on the device with
no host synchronisation inside.  The timed window sits in the middle of the
run: untimed spin-up steps, W untimed warmup steps, then exactly K timed
steps (the default W=0, K=5000 times the whole run).

Metric: cell-stage updates/s = 2 * (active cells summed over the timed steps)
/ wall time, where active cells are the check_solve set of each step (the
reference's definition, which includes the always-active boundary frame);
config.work_cell_stage_updates_per_s counts only the cells the update kernel
processes.  config.whole_run repeats the workload's whole run in the same
process (wall clock, no events).

The JSON line also carries
  roofline:     the dominant kernel's algorithmic HBM bytes per launch (per
                processed cell for the update kernel) over its mean launch
                time from HIP events on the launch stream -- around the
                launches of every E-th timed step for K >= 500
                (--event-every), else around every launch of the K steps
                that follow the timed window (events would perturb a short
                window) -- vs 8 TB/s; traffic = measured HBM bytes per
                launch when the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
                summary covers those same steps, else null;
  cpu_baseline: the reference OpenMP Fortran build (compiled from its own
                sources into oracle/_ref) timed on this host's physical
                cores over a bounded sample of steps that starts from the
                GPU state at the middle of the timed region.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

HBM_PEAK_GBS = 8000.0


# the three kernels of a device-side driver step, in launch order, as
# imex_run_steps_timed times them (rocprofv3 names in brackets)
KERNELS = ["k_mask", "k_flux", "k_update"]

# driver steps of each workload's run (the timed window is placed in its middle)
RUN_STEPS = {"c1": 1000, "c2": 5000, "c2-small": 1000}


def kernel_bytes_per_cell(cfg):
    """Algorithmic HBM bytes per active cell per launch of each kernel: the
    whole-step figure 8 (3 n_v + 2) + 16 n_s S + 57 split over the kernels.

    k_mask [k_mask]:          read q(1), write the mask byte        9
    k_flux [k_flux<NS>]:      read q (n_v), bed slopes B'_x, B'_y   8 (n_v + 2)
    k_update [k_update<NS,7>]: read and write q (16 n_v), deposit
                              read and write with settling
                              (16 n_s), hmax / pdynmax /
                              mod_vel_max read and write (48)       16 n_v + 16 n_s S + 48
    Intermediates that a fused single pass would not need (the flux
    divergence handed from k_flux to k_update, the tile and work lists) are
    not counted: they show up in the measured traffic instead.
    """
    nv, ns = cfg.n_vars, cfg.n_solid
    s = 1 if cfg.settling_flag else 0
    per = {"k_mask": 9, "k_flux": 8 * (nv + 2), "k_update": 16 * nv + 16 * ns * s + 48}
    per["step"] = sum(per.values())
    return per


def read_traffic(kernel: str, workload: str, window):
    """Measured HBM bytes per processed cell of one kernel of the step, from
    the committed rocprofv3 PMC summary profiles/<workload>_step_kernels_traffic.md
    (lines 'window: steps A..B' and 'bytes_per_cell <kernel>: X'), when its
    window is the timed window (first, last) of this run; else None."""
    p = os.path.join(HERE, "profiles", f"{workload}_step_kernels_traffic.md")
    if not os.path.exists(p):
        return None
    text = open(p).read()
    w = re.search(r"^window: steps (\d+)\.\.(\d+)", text, re.M)
    if not w or (int(w.group(1)), int(w.group(2))) != tuple(window):
        return None
    m = re.search(rf"^bytes_per_cell {re.escape(kernel)}:\s*([0-9.eE+-]+)", text, re.M)
    return float(m.group(1)) if m else None


def host_cores(share: int = 16):
    """Threads for the CPU baseline: the physical cores (lscpu's CORE,SOCKET
    pairs) among the CPUs this process may run on, at most the host's CPU
    share of one GPU; and lscpu's topology line."""
    cpus = sorted(os.sched_getaffinity(0))
    topo, core_of = "", {}
    try:
        out = subprocess.run(["lscpu", "-p=CPU,CORE,SOCKET"], capture_output=True, text=True,
                             timeout=30).stdout
        for ln in out.splitlines():
            if ln and not ln.startswith("#"):
                c, core, sock = (int(x) for x in ln.split(",")[:3])
                core_of[c] = (sock, core)
        info = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
        def get(key):
            m = re.search(r"^" + re.escape(key) + r":\s*(.+)$", info, re.M)
            return m.group(1).strip() if m else "?"
        topo = (get("Socket(s)") + " socket(s) x " + get("Core(s) per socket") + " cores x "
                + get("Thread(s) per core") + " threads, " + get("Model name"))
    except (OSError, subprocess.SubprocessError):
        pass
    phys = {core_of[c] for c in cpus if c in core_of} or set(cpus)
    return max(1, min(len(phys), share)), topo


def cpu_baseline(workload: str, state: dict, seconds: float, threads: int = 0,
                 max_steps: int = 200):
    """Runs oracle/ref_timing.py in a child process on a bounded sample, with
    OMP_NUM_THREADS = the physical cores available (host_cores), bound
    close to one thread per core."""
    from oracle import refpy
    if not refpy.available("perf"):
        return None
    n, topo = host_cores()
    if threads:
        n = threads
    fd, path = tempfile.mkstemp(suffix=".npz", prefix="imex_state_")
    os.close(fd)
    try:
        np.savez(path, **state)
        env = dict(os.environ)
        env.update(OMP_NUM_THREADS=str(n), OMP_PROC_BIND="close", OMP_PLACES="cores")
        env.setdefault("OMP_STACKSIZE", "1G")
        cmd = [sys.executable, "-m", "oracle.ref_timing", "--workload", workload,
               "--state", path, "--seconds", str(seconds), "--max-steps", str(max_steps)]
        # the child's progress lines go to this process's stderr
        r = subprocess.run(cmd, cwd=HERE, env=env, stdin=subprocess.DEVNULL,
                           stdout=subprocess.PIPE, stderr=None, text=True, timeout=3000)
        if r.returncode != 0:
            print(f"cpu_baseline: reference run failed ({r.returncode})", file=sys.stderr,
                  flush=True)
            return None
        line = [ln for ln in r.stdout.splitlines() if ln.startswith("{")][-1]
        res = json.loads(line)
    finally:
        os.unlink(path)
    rate = 2.0 * res["active_cells"] / res["elapsed_s"]
    return {
        "value": rate, "unit": "cell-stage updates/s", "cores": res["threads"],
        "kind": "reference",
        "sample": (f"{res['steps']} driver steps of the same run from its state at step "
                   f"{state['step']}, {res['active_cells']} active cells in total, "
                   f"{res['elapsed_s']:.1f} s; reference src/*.f90 built with "
                   f"amdflang -O3 -ffast-math -fstack-arrays -funroll-loops -fopenmp (flang's "
                   f"spelling of -Ofast; it has no -funroll-all-loops), "
                   f"OMP_NUM_THREADS={res['threads']} OMP_PROC_BIND=close OMP_PLACES=cores; "
                   f"host {topo}"),
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed driver steps (default: the workload's whole run less the warmup; "
                         "c3: 3000)")
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", default="c2",
                    help="c2 (default), c1, c2-small, c3, c4 (the batched ensemble), c5 "
                         "(serving), c2-outputs (the run with periodic outputs) or dambreak "
                         "(the reference's TEST_STATIC_FINH deck: pore pressure)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the reference CPU sample")
    ap.add_argument("--skip-whole-run", action="store_true",
                    help="do not repeat the workload's whole run after the timed window (for "
                         "counter passes, which then see exactly the spin-up, warmup and timed "
                         "launches)")
    ap.add_argument("--cpu-full", action="store_true",
                    help="also time the reference's whole run of the workload from its initial "
                         "state on all physical cores and on one thread (long: minutes)")
    ap.add_argument("--event-every", type=int, default=50, metavar="E",
                    help="HIP events around the kernel launches of every E-th timed step "
                         "(an event record costs the stream a few microseconds)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="N",
                    help="only run the plain copy kernel over N doubles 20 times (the "
                         "reference point for rocprofv3 FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--members", type=int, default=256, help="c4: ensemble members")
    ap.add_argument("--t-end", type=float, default=30.0, help="c4: physical end time of a member (s)")
    ap.add_argument("--rates", default="2,8,32", help="c5: arrival rates per second")
    ap.add_argument("--window", type=float, default=20.0, help="c5: arrival window (wall s)")
    ap.add_argument("--serve-t-end", default="60,600", help="c5: t_end range of a request (s)")
    ap.add_argument("--serve-classes", type=int, default=1,
                    help="c5: solid classes of every request (1: the ensemble members' physics; "
                         "2 or 3 with settling)")
    ap.add_argument("--out-t-end", type=float, default=200.0,
                    help="c2-outputs: physical end time (s)")
    ap.add_argument("--out-dt", type=float, default=100.0,
                    help="c2-outputs: interval of the solution outputs (s)")
    ap.add_argument("--info", default=None,
                    help="also write per-kernel launch counts and active cells to this JSON file")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the multi-GPU decomposition is a design, not part of this build)",
              file=sys.stderr)
        sys.exit(2)
    args.steps_given = args.steps is not None
    if args.steps is None:
        args.steps = max(1, RUN_STEPS.get(args.workload, 5000) - args.warmup)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be >= 1 and --warmup >= 0")

    from imex_sflow2d_dev_amd import init, native, workloads
    from imex_sflow2d_dev_amd.solver import Solver
    if not os.path.exists(native.LIB_PATH):
        from imex_sflow2d_dev_amd import build
        build.build()

    if args.calib_copy:
        h = native.Handle(workloads.by_name("c2-small")[0])
        h.check(h.L.imex_calib_copy(h.h, ctypes.c_int64(args.calib_copy), 20), "imex_calib_copy")
        h.close()
        print(json.dumps({"calib_copy_doubles": args.calib_copy, "launches": 20,
                          "bytes_read_per_launch": 8 * args.calib_copy,
                          "bytes_written_per_launch": 8 * args.calib_copy}), flush=True)
        return

    if args.workload == "c4":
        bench_c4(args)
        return
    if args.workload == "c5":
        bench_c5(args)
        return
    if args.workload == "c3":
        bench_c3(args)
        return
    if args.workload == "c2-outputs":
        bench_outputs(args)
        return
    if args.workload == "dambreak":
        bench_dambreak(args)
        return

    t_host = time.perf_counter()
    cfg, B = workloads.by_name(args.workload)
    static, q0 = init.setup(cfg, B)
    full = RUN_STEPS.get(args.workload, args.warmup + args.steps)
    # the timed window sits in the middle of the workload's run (the flow of
    # its first and last steps is not representative of the run)
    spinup = max(0, (full - args.warmup - args.steps) // 2)
    print(f"[bench] workload {args.workload}: {cfg.nx}x{cfg.ny} n_solid={cfg.n_solid} "
          f"host set-up {time.perf_counter() - t_host:.1f} s; run of {full} steps, "
          f"{spinup} spin-up steps before the warmup", file=sys.stderr, flush=True)

    # load code objects and touch every kernel once on a throw-away run
    s = Solver(cfg, static, q0)
    s.run_steps(3)
    s.close()

    t_up = time.perf_counter()                  # handle, static fields and q0 from host memory
    s = Solver(cfg, static, q0)
    s_time(s)
    t_up = time.perf_counter() - t_up
    s.raise_on_error = False                    # physics error bits are reported, not raised
    if spinup:
        s.run_steps(spinup, record_dt=False)
    if args.warmup:
        s.run_steps_timed(args.warmup, record_dt=False, sample_every=args.warmup)
    print(f"[bench] spin-up {spinup} and warmup {args.warmup} steps done, t={s.t:.4f}",
          file=sys.stderr, flush=True)
    first = spinup + args.warmup + 1            # driver step numbers of the timed window
    last = spinup + args.warmup + args.steps

    # timed region: two halves, with the mid-run state copied out between
    # them (outside the clock) for the CPU sample
    halves = [args.steps // 2, args.steps - args.steps // 2]
    kernel_ms = np.zeros(len(KERNELS))
    sampled = 0
    cells = 0
    wall = 0.0
    mid = None
    err_bits = 0
    work0, frame = s.work_counters()
    work = 0
    # A short window carries no HIP events (each event record costs the
    # stream microseconds, a sizeable share of a 20-step window): its kernel
    # durations come from the same number of steps right after it, with
    # events around every launch.  A long window samples every E-th step.
    events_inside = args.steps >= 500
    for i, n in enumerate(halves):
        if n == 0:
            continue
        s_time(s)                               # synchronises the stream
        a0, _ = s.counters()
        w0 = s.work_counters()[0]
        t0 = time.perf_counter()
        # both return after a stream sync
        if events_inside:
            _, k, ns, c = s.run_steps_timed(n, sample_every=args.event_every)
        else:
            s.run_steps(n, record_dt=False)
            k, ns = np.zeros(len(KERNELS)), 0
        t1 = time.perf_counter()
        c = s.counters()[0] - a0
        work += s.work_counters()[0] - w0
        err_bits |= s.err
        wall += t1 - t0
        kernel_ms += k
        sampled += ns
        cells += c
        print(f"[bench] timed part {i + 1}: {n} steps, {t1 - t0:.3f} s, "
              f"{c} active cells", file=sys.stderr, flush=True)
        if i == 0:
            mid = dict(q=s.q, t=s.t, dt_old=s.dt_old, dt_old_old=s.dt_old_old,
                       step=first - 1 + n)
    if mid is None:
        mid = dict(q=s.q, t=s.t, dt_old=s.dt_old, dt_old_old=s.dt_old_old, step=last)
    if events_inside:
        kwin = (first, last)
        kcells, kwork = cells, work
    else:
        a0, _ = s.counters()
        w0 = s.work_counters()[0]
        _, kernel_ms, sampled, _ = s.run_steps_timed(args.steps, record_dt=False, sample_every=1)
        kcells = s.counters()[0] - a0
        kwork = s.work_counters()[0] - w0
        err_bits |= s.err
        kwin = (last + 1, last + args.steps)
    t_end_phys = s.t
    t_down = time.perf_counter()                # the state and the output fields to host memory
    s.q
    s.fields()
    t_down = time.perf_counter() - t_down
    s.close()

    K = args.steps
    value = 2.0 * cells / wall
    # the whole run of the workload in the same process (wall clock, no events)
    if args.skip_whole_run:
        whole = dict(steps=K, wall_s=wall, active_cells=cells, work_cells=work,
                     physics_error_bits=err_bits, note="timed window only (--skip-whole-run)")
    elif (spinup, args.warmup, K) == (0, 0, full):
        whole = dict(steps=K, wall_s=wall, active_cells=cells, work_cells=work,
                     physics_error_bits=err_bits)
    else:
        s2 = Solver(cfg, static, q0)
        s2.raise_on_error = False
        s_time(s2)
        a_init, _ = s2.counters()               # the upload's check_solve
        t0 = time.perf_counter()
        s2.run_steps(full, record_dt=False)
        s_time(s2)
        wall2 = time.perf_counter() - t0
        a2, _ = s2.counters()
        w2, _ = s2.work_counters()
        whole = dict(steps=full, wall_s=wall2, active_cells=a2 - a_init, work_cells=w2,
                     physics_error_bits=s2.err)
        s2.close()
    whole["ms_per_step"] = whole["wall_s"] * 1e3 / whole["steps"]
    whole["cell_stage_updates_per_s"] = 2.0 * whole["active_cells"] / whole["wall_s"]
    whole["work_cell_stage_updates_per_s"] = 2.0 * whole["work_cells"] / whole["wall_s"]

    per = kernel_bytes_per_cell(cfg)
    names = KERNELS
    dom = int(np.argmax(kernel_ms))
    dname = names[dom]
    mean_launch_s = kernel_ms[dom] * 1e-3 / sampled
    klaunches = kwin[1] - kwin[0] + 1
    cells_per_launch = kcells / klaunches
    work_per_launch = kwork / klaunches
    # the update kernel visits only its work list; the others are measured on
    # the active cells (their per-cell bytes are those of an active cell)
    basis_cells = work_per_launch if dname == "k_update" else cells_per_launch
    alg_bytes = per[dname] * basis_cells
    achieved = alg_bytes / mean_launch_s / 1e9
    achieved_active = per[dname] * cells_per_launch / mean_launch_s / 1e9
    tb = read_traffic(dname, args.workload, kwin)
    roofline = {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS,
        "traffic": (tb * basis_cells) if tb is not None else None,
        "kernel": dname,
        "cells_basis": "work cells (the cells the kernel processes)" if dname == "k_update"
                       else "active cells",
        "alg_bytes_per_cell": per[dname],
        "cells_per_launch": basis_cells,
        "achieved_on_active_cells": achieved_active,
        "mean_launch_us": mean_launch_s * 1e6,
        "event_timed_launches": sampled,
        "kernel_window_steps": list(kwin),
        "mean_active_cells_per_launch": cells_per_launch,
    }
    step_gbs = per["step"] * cells / wall / 1e9
    out = {
        "metric": "cell-stage updates/s",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 DEM and source, generated in-process)",
        "config": {
            "workload": (f"{args.workload}: {cfg.nx}x{cfg.ny} cells dx={cfg.dx} m, "
                         f"{cfg.n_solid} solid classes, settling={cfg.settling_flag}, "
                         f"radial source={cfg.radial_source_flag}, FP64, timed driver steps "
                         f"{first}..{last} of the {full}-step run"),
            "grid": [cfg.nx, cfg.ny],
            "run_steps": full,
            "spinup_steps": spinup,
            "timed_steps": [first, last],
            "active_cells_timed": cells,
            "work_cells_timed": work,
            "frame_cells_per_step": frame,
            "work_cell_stage_updates_per_s": 2.0 * work / wall,
            "active_fraction": cells / (K * cfg.nx * cfg.ny),
            "grid_cell_stage_updates_per_s": 2.0 * cfg.nx * cfg.ny * K / wall,
            "t_physical_end_s": t_end_phys,
            "step_hbm_alg_GBs": step_gbs,
            "step_alg_bytes_per_active_cell": per["step"],
            "kernel_ms_event_timed_steps": dict(zip(names, [float(x) for x in kernel_ms])),
            "event_timed_steps": sampled,
            "physics_error_bits": err_bits,
            "whole_run": whole,
            "host_upload_s": t_up,
            "host_download_s": t_down,
            "host_buffers_inclusive_cell_stage_updates_per_s": 2.0 * cells / (wall + t_up + t_down),
        },
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if args.info:
        with open(args.info, "w") as f:
            json.dump({"spinup": spinup, "timed_steps": [first, last],
                       "active_cells_timed": cells, "work_cells_timed": work,
                       "kernel_ms_event_timed_steps": out["config"]["kernel_ms_event_timed_steps"],
                       "event_timed_steps": sampled,
                       "alg_bytes": per}, f, indent=1)
    if not args.no_cpu:
        print("[bench] CPU reference sample ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload, mid, args.cpu_seconds)
    if args.cpu_full:
        init_state = dict(q=q0, t=cfg.t_start, dt_old=cfg.dt0, dt_old_old=cfg.dt0, step=0)
        runs = []
        for th in (host_cores()[0], 1):
            print(f"[bench] reference whole run on {th} thread(s) ...", file=sys.stderr,
                  flush=True)
            runs.append(cpu_baseline(args.workload, init_state, 1e9, threads=th,
                                     max_steps=full))
        out["config"]["cpu_whole_run"] = runs
    print(json.dumps(out), flush=True)


def ensemble_host_baseline(r, t_end: float, seconds: float = 20.0):
    """The reference's own ensemble pattern on this host
    (EXAMPLES/EXAMPLE_ENSEMBLE_ETNA/launch_jobs.py: independent runs, one
    single-threaded process each): a sample of members spread over the
    ensemble's grid sizes, one process per physical core at once, each timed
    for up to `seconds` on the reference build (oracle/ref_ensemble.py).  A
    member that does not finish is extrapolated by its steps: time = elapsed
    x (the member's step count in the batch) / (steps done) -- its early,
    smaller flow makes this an underestimate of the CPU time.  runs/hr = the
    ensemble's member count over (mean member time x members / cores)."""
    from oracle import refpy
    if not refpy.available("perf"):
        return None
    n, topo = host_cores()
    members = sorted(set(np.linspace(0, r["members"] - 1, n).astype(int).tolist()))
    env = dict(os.environ)
    env.update(OMP_NUM_THREADS="1", OMP_PROC_BIND="true")
    env.setdefault("OMP_STACKSIZE", "256M")
    procs = []
    for m in members:
        cmd = [sys.executable, "-m", "oracle.ref_ensemble", "--member", str(m),
               "--t-end", str(t_end), "--seconds", str(seconds)]
        procs.append((m, subprocess.Popen(cmd, cwd=HERE, env=env, stdin=subprocess.DEVNULL,
                                          stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                          text=True)))
    res = []
    for m, p in procs:
        out, err = p.communicate(timeout=seconds * 20 + 600)
        if p.returncode != 0:
            print(f"ensemble baseline: member {m} failed ({p.returncode}): {err[-1000:]}",
                  file=sys.stderr, flush=True)
            return None
        d = json.loads([ln for ln in out.splitlines() if ln.startswith("{")][-1])
        need = r["_member_steps"][m]
        d["member_time_s"] = (d["elapsed_s"] if d["finished"]
                              else d["elapsed_s"] * need / max(d["steps"], 1))
        res.append(d)
    mean_t = float(np.mean([d["member_time_s"] for d in res]))
    wall = mean_t * r["members"] / len(members)
    return {
        "value": r["members"] / wall * 3600.0, "unit": "runs/hr", "cores": len(members),
        "kind": "reference",
        "sample": (f"members {members} of the {r['members']}, each a single-threaded process of "
                   f"the reference build (OMP_NUM_THREADS=1, {len(members)} at once on "
                   f"{len(members)} physical cores), timed for up to {seconds:.0f} s each; "
                   f"{sum(d['finished'] for d in res)} finished, the others extrapolated by "
                   f"their batch step counts; mean member time {mean_t:.1f} s; host {topo}"),
    }


def bench_c4(args):
    """The batched ensemble (256 members of 128^2..1024^2 cells, each to its
    own t_end) against the same runs one after the other on the same GPU, in
    FP64 (value) and in the FP32 form (beside it); the reference's own
    ensemble pattern on the host cores as cpu_baseline."""
    from imex_sflow2d_dev_amd import ensemble
    say = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    r = ensemble.bench_ensemble(n_members=args.members, t_end=args.t_end, log=say)
    r32 = ensemble.bench_ensemble(n_members=args.members, t_end=args.t_end, log=say, precision=1)
    base = None if args.no_cpu else ensemble_host_baseline(r, args.t_end)
    r.pop("_member_steps")
    r32.pop("_member_steps")
    out = {
        "metric": "ensemble runs/hr",
        "value": r["runs_per_hour_batched"],
        "unit": "runs/hr",
        "n_gpus": 1,
        "steps": r["member_steps_total"],
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 fractal DEMs and member parameters, generated in-process)",
        "config": dict({"workload": (f"c4: {r['members']} members, sides log-uniform in 128..1024 "
                                     f"cells (dx 10 m), each to t_end = {r['t_end_s']} s, FP64; "
                                     f"the FP32 form beside it")}, **r, fp32_form=r32),
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


def bench_c3(args):
    """The precision form: the 4096^2, three-class run for --steps steps
    (3000 by default here) in FP64 and in the FP32 form; value = the FP32
    form's cell-stage updates/s, beside the FP64 rate and the relative
    differences of runout area, maximum thickness, deposit and mass."""
    from imex_sflow2d_dev_amd import init, precision, workloads
    steps = args.steps if args.steps_given else 3000
    say = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    cfg, B = workloads.config3()
    static, q0 = init.setup(cfg, B)
    r = precision.compare(cfg, static, q0, steps, log=say)
    # the same run with sedimentation on, for the deposit (the 1 cm class
    # settles within a few hundred steps, so this variant is not timed)
    cfg_s, _ = workloads.config3()
    cfg_s.settling_flag = True
    rs = precision.compare(cfg_s, static, q0, steps, log=say)
    out = {
        "metric": "cell-stage updates/s (FP32 form; FP64 beside it)",
        "value": r["fp32"]["cell_stage_updates_per_s"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": r["fp32"]["wall_s"] * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (state, reconstruction, fluxes) / f64 (Newton, thin layers, deposit, budget)",
        "data": "synthetic (seeded PCG64 DEM, generated in-process)",
        "config": {"workload": (f"c3: {cfg.nx}x{cfg.ny} cells dx={cfg.dx} m, {cfg.n_solid} solid classes, "
                                f"settling={cfg.settling_flag}, driver steps 1..{steps}"),
                   "fp64": r["fp64"], "fp32": r["fp32"],
                   "relative_difference_fp32_vs_fp64": r["relative_difference"],
                   "speedup_fp32": r["speedup_fp32"],
                   "with_sedimentation": {"fp64": rs["fp64"], "fp32": rs["fp32"],
                                          "relative_difference_fp32_vs_fp64":
                                              rs["relative_difference"]}},
    }
    print(json.dumps(out), flush=True)


def bench_dambreak(args):
    """The reference's TEST_STATIC_FINH deck (TESTS/TEST_DAMBREAK: pore
    pressure with f_inhibit STATIC, mu(I) friction, gas loss, the
    per-variable boundaries; 300 x 15 cells of 1 cm from its restart state to
    T_END = 0.3 s) on the general path (csrc/imex_pore.hip), device-driven;
    value = cell-stage updates/s of the whole run (2 x active cells summed
    over the steps / wall time), timed from the uploaded state after an
    untimed run that finds the step count.  Beside it, the reference's own
    run of the deck (oracle/_ref, -Ofast build, one OpenMP thread: with
    several its cell loops race on the module variable
    hydraulic_permeability)."""
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests"))
    import cases  # the deck's configuration and restart state
    import numpy as np
    from imex_sflow2d_dev_amd.solver import Solver
    say = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    cfg, B, st, q0 = cases.dambreak_deck("STATIC")
    s = Solver(cfg, st, q0, raise_on_error=False)
    steps = int(np.count_nonzero(s.run_steps(20000, t_end=cfg.t_end)))
    s.close()
    say(f"dambreak STATIC: {steps} steps to t_end {cfg.t_end}")
    s = Solver(cfg, st, q0, raise_on_error=False)
    a0, _ = s.counters()
    t0 = time.perf_counter()
    dts = s.run_steps(steps, t_end=cfg.t_end)
    wall = time.perf_counter() - t0
    a1, _ = s.counters()
    err = s.err
    s.close()
    value = 2.0 * (a1 - a0) / wall
    out = {
        "metric": "cell-stage updates/s",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "the reference deck's restart state (tests/golden/dambreak_q0.npz)",
        "config": {"workload": (f"dambreak: TEST_STATIC_FINH, {cfg.nx}x{cfg.ny} cells dx={cfg.dx} m, "
                                f"pore pressure, f_inhibit STATIC, to t_end {cfg.t_end} s"),
                   "active_cells": int(a1 - a0), "t_end_reached": float(np.sum(dts)),
                   "physics_error_bits": int(err), "wall_s": wall},
    }
    if not args.no_cpu:
        from oracle import refpy
        if refpy.available("perf"):
            refpy.set_threads(1, "perf")
            run = refpy.RefRun(cfg, B, flavour="perf")
            run.set_q(q0)
            t0 = time.perf_counter()
            _, ncells = run.run(steps)
            cw = time.perf_counter() - t0
            run.close()
            out["cpu_baseline"] = {"value": 2.0 * float(np.sum(ncells)) / cw,
                                   "unit": "cell-stage updates/s",
                                   "cores": 1, "kind": "reference",
                                   "sample": f"the whole deck, {steps} driver steps, {cw:.2f} s"}
    print(json.dumps(out), flush=True)


def bench_c5(args):
    """Serving: Poisson arrivals at each rate of --rates over --window seconds
    of wall time, admitted by device-memory footprint into one running batch;
    time-to-result percentiles and cell-stage updates/s beside a FIFO server
    of plain runs."""
    from imex_sflow2d_dev_amd import serving
    t_end = tuple(float(x) for x in args.serve_t_end.split(","))
    res = []
    for rate in [float(x) for x in args.rates.split(",")]:
        r = serving.bench_serving(rate, args.window, t_end=t_end, n_solid=args.serve_classes,
                                  log=lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True))
        print(f"[bench] rate {rate}/s: {json.dumps(r)}", file=sys.stderr, flush=True)
        res.append(r)
    top = res[-1]
    out = {
        "metric": "p50/p99 time-to-result (s)",
        "value": top["p99_s"],
        "unit": "s (p99 time-to-result at the highest rate)",
        "n_gpus": 1,
        "steps": sum(r["steps_total"] for r in res),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 request stream, DEMs and parameters, generated in-process)",
        "config": {"workload": (f"c5: Poisson arrivals at {args.rates} per s over {args.window} s, sides "
                                f"log-uniform in 64..2048 cells (dx 10 m), t_end U[{args.serve_t_end}] s, "
                                f"200 GB admission budget, {args.serve_classes} solid class(es), FP64"),
                   "rates": res},
    }
    print(json.dumps(out), flush=True)


def bench_outputs(args):
    """The radial-source flow at 2048^2 (c2) to t = --out-t-end physical
    seconds with the reference's periodic outputs on the device state:
    output_runout every 10 s, 16 probes every 5 s, output_solution (.q_2d,
    deposit .q_2d, .p_2d) and restart.bin every --out-dt s, against the same
    run without outputs.  value = cell-stage updates/s with the outputs;
    per-output costs and the writers' throughput beside it.  cpu_baseline:
    the Fortran runtime writing the same .p_2d records single-threaded
    (oracle/_ref/records_writer, the reference's own format statement) and
    the whole-array runout expressions in numpy, on a sample."""
    import shutil
    from imex_sflow2d_dev_amd import init, native, outputs, workloads
    from imex_sflow2d_dev_amd.solver import Solver
    say = lambda m: print(f"[bench] {m}", file=sys.stderr, flush=True)  # noqa: E731
    cfg, B = workloads.by_name("c2")
    static, q0 = init.setup(cfg, B)
    t_end, dt_out = args.out_t_end, args.out_dt
    # the file system with the most free space; one solution output of this
    # grid writes about 3.3 GB
    cands = [d for d in ("/dev/shm", tempfile.gettempdir()) if os.path.isdir(d)]
    base = max(cands, key=lambda d: shutil.disk_usage(d).free)
    free = shutil.disk_usage(base).free
    if 3.5e9 * (int(t_end // dt_out) + 1) > 0.8 * free:
        dt_out = t_end
    if 3.5e9 * 2 > 0.8 * free:
        sys.exit(f"c2-outputs needs about 7 GB free for its files; {base} has {free / 1e9:.1f} GB")
    say(f"files in {base} ({free / 1e9:.0f} GB free), solution outputs every {dt_out} s")
    work = tempfile.mkdtemp(prefix="imex_outputs_", dir=base)
    res = {}
    try:
        # the same run without outputs
        s = Solver(cfg, static, q0, raise_on_error=False)
        s.run_steps(3, t_end=cfg.t_start)      # loads the code objects, takes no step
        a0, _ = s.counters()
        t0 = time.perf_counter()
        n_plain = 0
        while s.t < t_end:
            d = s.run_steps(256, t_end=t_end)
            n_plain += int(np.count_nonzero(d))
        wall_plain = time.perf_counter() - t0
        cells_plain = s.counters()[0] - a0
        s.close()
        say(f"no outputs: {n_plain} steps to t={t_end} in {wall_plain:.3f} s")

        s = Solver(cfg, static, q0, raise_on_error=False)
        s.run_steps(3, t_end=cfg.t_start)
        xs, ys = cfg.x_source, cfg.y_source
        ang = np.linspace(0.0, 2.0 * np.pi, 16, endpoint=False)
        coords = np.stack([xs + 800.0 * np.cos(ang), ys + 800.0 * np.sin(ang)], 1)
        ro = outputs.RunoutOutput(s, "c2", work, dt_runout=10.0, B=B)
        po = outputs.ProbesOutput(s, coords, "c2", work, dt_probes=5.0)
        so = outputs.SolutionOutput(s, "c2", work, dt_output=dt_out, B=B, n_threads=16)
        calls = {"runout": [], "probes": []}

        def timed(fn, key):
            def call(*a):
                t = time.perf_counter()
                r = fn(*a)
                calls[key].append(time.perf_counter() - t)
                return r
            return call

        class W:  # a writer's calls timed, its state shared
            def __init__(self, inner, key):
                object.__setattr__(self, "_i", inner)
                object.__setattr__(self, "_c", timed(inner.__call__, key))

            def __call__(self, *a):
                return self._c(*a)

            def __getattr__(self, n):
                return getattr(self._i, n)

            def __setattr__(self, n, v):
                setattr(self._i, n, v)

        a0, _ = s.counters()
        t0 = time.perf_counter()
        steps, _ = outputs.run_with_outputs(s, t_end, runout=W(ro, "runout"), probes=W(po, "probes"),
                                            solution=so, chunk=256)
        wall_out = time.perf_counter() - t0
        cells_out = s.counters()[0] - a0
        sizes = {f: os.path.getsize(os.path.join(work, f)) for f in os.listdir(work)}
        sol_bytes = sum(v for f, v in sizes.items() if f.endswith((".q_2d", ".p_2d", ".bin")))
        say(f"with outputs: {steps} steps in {wall_out:.3f} s; {len(calls['runout'])} runout, "
            f"{len(calls['probes'])} probe, {so.output_idx} solution outputs")
        # device costs of one call, on the final state
        t = time.perf_counter()
        for _ in range(20):
            s.runout_reduce(ro.x0, ro.y0, 0.0)
        red_s = (time.perf_counter() - t) / 20
        t = time.perf_counter()
        for _ in range(20):
            s.probes(coords)
        probe_s = (time.perf_counter() - t) / 20
        N = cfg.nx * cfg.ny
        p2d = so.bufs[native.SNAP_P2D].a.reshape(N, -1)
        # writer throughput on one .p_2d file: 16 threads and 1 thread
        wt = {}
        for nt in (16, 1):
            path = os.path.join(work, f"tp_{nt}.p_2d")
            t = time.perf_counter()
            native.write_records(path, p2d, cfg.nx, native.FMT_ES15_7, nt)
            wt[nt] = time.perf_counter() - t
            p2d_bytes = os.path.getsize(path)
            os.remove(path)
        res = {
            "wall_s": wall_out, "steps": steps, "cells": cells_out, "physics_error_bits": s.err,
            "cell_stage_updates_per_s": 2.0 * cells_out / wall_out,
            "no_outputs": {"wall_s": wall_plain, "steps": n_plain,
                           "cell_stage_updates_per_s": 2.0 * cells_plain / wall_plain},
            "outputs": {"runout_calls": len(calls["runout"]),
                        "runout_call_ms_mean": 1e3 * float(np.mean(calls["runout"])),
                        "runout_reduce_ms": 1e3 * red_s,
                        "probe_calls": len(calls["probes"]),
                        "probe_call_ms_mean": 1e3 * float(np.mean(calls["probes"])),
                        "probes_device_ms": 1e3 * probe_s, "n_probes": len(coords),
                        "solution_outputs": so.output_idx,
                        "solution_caller_ms": [1e3 * a for a, _ in so.timings],
                        "solution_writer_s": [b for _, b in so.timings],
                        "solution_bytes_total": sol_bytes,
                        "files": len(sizes)},
            "p2d_file_bytes": p2d_bytes,
            "p2d_write_s": {"16_threads": wt[16], "1_thread": wt[1]},
            "p2d_write_MBps_16_threads": p2d_bytes / wt[16] / 1e6,
        }
        sample = np.array(p2d[:128 * cfg.nx])     # the page-locked buffers go with so.close()
        so.close()
        if not args.no_cpu:
            res["cpu_baseline"] = outputs_cpu_baseline(cfg, s, sample, work, say)
        s.close()
    finally:
        shutil.rmtree(work, ignore_errors=True)
    out = {
        "metric": "cell-stage updates/s with periodic outputs",
        "value": res["cell_stage_updates_per_s"],
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": res["steps"],
        "warmup": 0,
        "ms_per_step": res["wall_s"] * 1e3 / res["steps"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded PCG64 DEM and source, generated in-process)",
        "config": {"workload": (f"c2-outputs: {cfg.nx}x{cfg.ny} cells, radial source, to t = {t_end} s "
                                f"with output_runout every 10 s, 16 probes every 5 s, output_solution "
                                f"(.q_2d, deposit .q_2d, .p_2d) and restart.bin every {dt_out} s, "
                                f"files in {base}"),
                   **{k: v for k, v in res.items() if k != "cpu_baseline"}},
    }
    if "cpu_baseline" in res:
        out["cpu_baseline"] = res["cpu_baseline"]
    print(json.dumps(out), flush=True)


def outputs_cpu_baseline(cfg, s, p2d, work, say):
    """The Fortran runtime writing .p_2d records single-threaded (the
    reference's format statement, oracle/_ref/records_writer) on a sample
    of rows of cells (p2d), and the whole-array runout expressions of
    output_runout in numpy (oracle/outputs_ref.py) on the full grid."""
    from oracle import oraclepy, outputs_ref
    exe = os.path.join(HERE, "oracle", "_ref", "records_writer")
    out = {"kind": "port", "cores": 1}
    if os.path.exists(exe):
        rec = np.ascontiguousarray(p2d)
        rows = rec.shape[0] // cfg.nx
        raw = os.path.join(work, "sample.raw")
        rec.tofile(raw)
        dst = os.path.join(work, "sample.p_2d")
        t = time.perf_counter()
        subprocess.run([exe, "p2d", raw, dst, str(cfg.nx), str(rows), str(rec.shape[1])], check=True)
        wf = time.perf_counter() - t
        nb = os.path.getsize(dst)
        out["p2d_write_MBps"] = nb / wf / 1e6
        out["p2d_sample"] = f"{rows} rows of {cfg.nx} cells, {nb} bytes, {wf:.2f} s"
        say(f"Fortran runtime .p_2d writer: {nb / wf / 1e6:.1f} MB/s")
    q = s.q
    o = oraclepy.Oracle(cfg)
    t = time.perf_counter()
    qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
    outputs_ref.runout_red(cfg, q, qp, np.zeros((cfg.ny, cfg.nx), bool), cfg.x_source, cfg.y_source, 0.0)
    out["runout_numpy_s"] = time.perf_counter() - t
    out["value"] = out.get("p2d_write_MBps")
    out["unit"] = "MB/s (.p_2d text, one thread)"
    out["sample"] = out.get("p2d_sample", "")
    return out


def s_time(s):
    """(t, dt) of the device-side driver; synchronises the handle's stream."""
    t = ctypes.c_double()
    dt = ctypes.c_double()
    s.H.check(s.L.imex_get_time(s.h, ctypes.byref(t), ctypes.byref(dt)), "get_time")
    return t.value, dt.value


if __name__ == "__main__":
    main()
