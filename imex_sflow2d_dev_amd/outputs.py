"""Host side of the driver's periodic outputs over the device state: the
runout files of output_runout (src/inpout_2d.f90:7170-7427), the probe files
of output_probes (:6926-7155) and the output schedule of the driver loop
(src/IMEX_SfloW2D.f90:395-591).  The grid work runs on the device
(Solver.runout_reduce, Solver.probes); this module keeps the reference's
state between calls and writes its files with its edit descriptors.

Files (in `directory`): <run_name>_runout.csv, <run_name>_mass_center.csv,
runout.txt, <run_name>_prb_NNNN.csv -- the names and line formats of the
reference.
"""
from __future__ import annotations

import math
import os

import numpy as np

from .fortran_format import a_fmt, es_fmt, f_fmt


def _fmax(a, b):
    """Fortran MAX(a, b) as the reference build evaluates it."""
    return a if a > b else b


def _div(a, b):
    """IEEE division (a / 0 gives +-Inf or NaN, as in the reference)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float64(a) / np.float64(b))


def lettera(k: int) -> str:
    """The four-digit index of output file names (src/inpout_2d.f90:6893-6912)."""
    return f"{k:04d}"


class RunoutOutput:
    """output_runout: runout distance, flow area, barycentre and the steady
    stop test, once per call, from one device reduction."""

    def __init__(self, solver, run_name: str = "run", directory: str = ".", dt_runout: float = 60.0,
                 eps_stop: float = 0.0, alphas_threshold: float = 0.0,
                 x0_runout: float = -1.0, y0_runout: float = -1.0, B=None):
        self.s = solver
        c = solver.cfg
        self.cfg = c
        self.dir = directory
        self.run_name = run_name
        self.dt_runout = dt_runout
        self.eps_stop = eps_stop
        self.thr = alphas_threshold
        # the reference location when the input leaves it unset: the collapse
        # centre, or the source centre (src/inpout_2d.f90:4817-4838)
        if x0_runout == -1.0 and y0_runout == -1.0:
            if c.collapsing_volume_flag:
                x0_runout, y0_runout = c.x_collapse, c.y_collapse
            if c.radial_source_flag:
                x0_runout, y0_runout = c.x_source, c.y_source
        self.x0, self.y0 = x0_runout, y0_runout
        if B is not None:
            solver.set_topography(B)
        self.has_b = B is not None
        self.init_runout = self.init_runout_x = self.init_runout_y = 0.0
        self.xmc_old = self.ymc_old = 0.0
        self.runout_last = 0.0
        self.t_runout = None
        self.dx = c.dx
        xs = c.x0 + np.arange(c.nx + 1) * c.dx
        ys = c.y0 + np.arange(c.ny + 1) * c.dx
        self.x_comp = 0.5 * (xs[:-1] + xs[1:])
        self.y_comp = 0.5 * (ys[:-1] + ys[1:])
        self.f_runout = os.path.join(directory, f"{run_name}_runout.csv")
        self.f_mc = os.path.join(directory, f"{run_name}_mass_center.csv")
        self.f_txt = os.path.join(directory, "runout.txt")
        # read_param opens both files (STATUS='unknown'); the first call writes
        # their headers
        open(self.f_runout, "w").close()
        open(self.f_mc, "w").close()

    def _xy(self, idx: int):
        return float(self.x_comp[idx % self.cfg.nx]), float(self.y_comp[idx // self.cfg.nx])

    def __call__(self, time: float, stop_flag: bool) -> bool:
        c = self.cfg
        s = self.s
        first = time == c.t_start
        if first:
            with open(self.f_runout, "a") as f:
                f.write(a_fmt("Time (s),", 13) + a_fmt("Runout (m),", 13) + a_fmt(" Area (m2),", 13)
                        + a_fmt("X_runout (m),", 17) + a_fmt("Y_runout (m)", 17) + "\n")
            with open(self.f_mc, "a") as f:
                f.write("".join(a_fmt(t, 18) for t in (
                    "Time (s),", "Barycenter-x (m),", "Barycenter-y (m),", "Baricenter-dx (m),",
                    "Baricenter-dy (m)")) + "\n")
            if self.x0 == -1.0 and self.y0 == -1.0:
                if not self.has_b:
                    raise ValueError("the runout origin is unset: give B_cent")
                rb = s.runout_reduce(self.x0, self.y0, self.thr, store=False, with_b=True)
                self.x0, self.y0 = self._xy(rb["ib"])
                r = s.runout_reduce(self.x0, self.y0, self.thr, store=True)
                # the first call's dist arrays are those of the common part
                self.init_runout = r["dist_max"]
                self.init_runout_x = r["dist_at_imax_x"]
                self.init_runout_y = r["dist_at_imax_y"]
            else:
                r = s.runout_reduce(self.x0, self.y0, self.thr, store=True)
            if r["max_h"] == 0.0:
                xmc, ymc = 0.0, 0.0
                if c.collapsing_volume_flag:
                    xmc, ymc = c.x_collapse, c.y_collapse
                if c.radial_source_flag:
                    xmc, ymc = c.x_source, c.y_source
            else:
                xmc = _div(r["sum_xq1"], r["sum_q1"])
                ymc = _div(r["sum_yq1"], r["sum_q1"])
            self.xmc_old, self.ymc_old = xmc, ymc
            old_runout = 0.0
        else:
            r = s.runout_reduce(self.x0, self.y0, self.thr, store=True)
            if r["max_h"] == 0.0:
                xmc, ymc = self.xmc_old, self.ymc_old
            else:
                xmc = _div(r["sum_xq1"], r["sum_q1"])
                ymc = _div(r["sum_yq1"], r["sum_q1"])
        if time > c.t_start:
            with open(self.f_txt) as f:
                old_runout = float(f.readline()[18:30])
        runout_now = r["dist_max"] - self.init_runout
        with open(self.f_txt, "w") as f:
            f.write(a_fmt("maximum runout =", 18) + f_fmt(_fmax(old_runout, runout_now), 12, 3) + "\n")
            f.write(a_fmt("final runout =", 18) + f_fmt(runout_now, 12, 3) + "\n")
        area = self.dx * self.dx * r["n_hpos"]
        area_old = area if first else self.dx * self.dx * r["n_hpos_old"]
        self.runout_last = runout_now
        X, Y = self._xy(r["imax"])
        with open(self.f_runout, "a") as f:
            f.write(f_fmt(time, 12, 3) + "," + f_fmt(self.runout_last, 12, 3) + ","
                    + f_fmt(area, 14, 3) + "," + f_fmt(X, 14, 3) + "," + f_fmt(Y, 14, 3) + "\n")
        with open(self.f_mc, "a") as f:
            f.write(f_fmt(time, 17, 3) + "," + f_fmt(xmc, 17, 3) + "," + f_fmt(ymc, 17, 3) + ","
                    + f_fmt(r["dist_x_max"] - self.init_runout_x, 17, 3) + ","
                    + f_fmt(r["dist_y_max"] - self.init_runout_y, 17, 3) + "\n")
        if time > c.t_start:
            vel_mass_center = math.sqrt((self.xmc_old - xmc) ** 2 + (self.ymc_old - ymc) ** 2) \
                / self.dt_runout
            area_new_rel = _div(self.dx * self.dx * r["n_new"], r["n_hpos"])
            self.xmc_old, self.ymc_old = xmc, ymc
            if _fmax(vel_mass_center, area_new_rel / self.dt_runout) < self.eps_stop and not stop_flag:
                stop_flag = True
        self.t_runout = time + self.dt_runout
        return stop_flag


class ProbesOutput:
    """output_probes: one CSV file per probe, one line per call."""

    def __init__(self, solver, coords, run_name: str = "run", directory: str = ".",
                 dt_probes: float = 60.0):
        self.s = solver
        self.cfg = solver.cfg
        self.xy = np.ascontiguousarray(coords, dtype=np.float64).reshape(-1, 2)
        self.files = [os.path.join(directory, f"{run_name}_prb_{lettera(k + 1)}.csv")
                      for k in range(len(self.xy))]
        self.dt_probes = dt_probes
        self.t_probes = None

    def _header(self):
        c = self.cfg
        ns = c.n_solid
        names = ["time,", "h,", "rho,", "T,", "u,", "v,"] + [f"alfa_s({i + 1:02d})," for i in range(ns)]
        units = ["(s),", "(m),", "(kg/m3),", "(K),", "(m/s),", "(m/s),"] + ["(),"] * ns
        if c.settling_flag:
            names += [f"dep_s({i + 1:02d})," for i in range(ns)]
            units += ["(m),"] * ns
        names.append("Dyn.pres.")
        units.append("(Pa)")
        return "".join(a_fmt(x, 15) for x in names) + "\n" + "".join(a_fmt(x, 15) for x in units) + "\n"

    def __call__(self, time: float):
        vals = self.s.probes(self.xy)
        first = time == self.cfg.t_start
        for k, path in enumerate(self.files):
            with open(path, "w" if first else "a") as f:
                if first:
                    f.write(f_fmt(self.xy[k, 0], 14, 3) + "," + f_fmt(self.xy[k, 1], 14, 3) + "\n")
                    f.write(self._header())
                row = vals[k]
                f.write(es_fmt(time, 14, 7) + ","
                        + "".join(es_fmt(v, 14, 7) + "," for v in row[:-1])
                        + es_fmt(row[-1], 14, 7) + "\n")
        self.t_probes = time + self.dt_probes


class SolutionOutput:
    """output_solution's .q_2d, _deposit_ .q_2d and .p_2d files
    (src/inpout_2d.f90:5934-6423) and write_restart_file's restart.bin
    (:8037-8084).  The device assembles the records (Solver.snapshot_begin)
    and copies them into page-locked buffers on its own stream; a host
    thread waits for the copy and writes the files while the next steps
    run.  A new output first waits for the previous one's files."""

    def __init__(self, solver, run_name: str = "run", directory: str = ".", dt_output: float = 60.0,
                 cons: bool = True, phys: bool = True, restart: bool = True, B=None,
                 n_threads: int = 0, overlap: bool = True):
        from . import native
        self.native = native
        self.s = solver
        c = solver.cfg
        self.cfg = c
        self.dir, self.run_name = directory, run_name
        self.dt_output = dt_output
        self.cons, self.phys, self.restart = cons, phys, restart
        self.n_threads, self.overlap = n_threads, overlap
        self.B = None if B is None else np.ascontiguousarray(B, dtype=np.float64)
        if B is not None:
            solver.set_topography(self.B)
        self.output_idx = 0
        self.t_output = c.t_start + dt_output if dt_output > 0 else c.t_end
        N = c.nx * c.ny
        self.what = 0
        if cons:
            self.what |= native.SNAP_Q2D | (native.SNAP_DEP2D if c.settling_flag else 0)
            if c.erosion_coeff > 0.0:
                self.what |= native.SNAP_ERO2D | native.SNAP_ERD2D
        if phys:
            self.what |= native.SNAP_P2D
        if restart or (cons and c.topo_change_flag):
            self.what |= native.SNAP_FIELDS
        if restart:
            self.what |= native.SNAP_Q
        self.bufs = {}
        for kind in native.SNAP_KINDS:
            if self.what & kind:
                self.bufs[kind] = native.PinnedArray(N * solver.snapshot_cols(kind))
        self._thread = None
        self._error = None
        self.timings = []   # per output: (seconds in the caller, seconds of the writer thread)

    def _files(self, idx: str, time: float, dt: float, with_restart: bool):
        import time as _t
        nat, c, s = self.native, self.cfg, self.s
        t0 = _t.perf_counter()
        try:
            s.snapshot_wait()
            N, nx = c.nx * c.ny, c.nx
            if self.cons:
                rec = self.bufs[nat.SNAP_Q2D].a.reshape(N, -1)
                nat.write_records(os.path.join(self.dir, f"{self.run_name}_{idx}.q_2d"), rec, nx,
                                  nat.FMT_E20_12, self.n_threads)
                if c.topo_change_flag:       # _B_ file (src/inpout_2d.f90:6077-6082)
                    fld = self.bufs[nat.SNAP_FIELDS].a.reshape(3 * c.n_solid + 4, N)
                    xs = c.x0 + np.arange(c.nx + 1) * c.dx
                    ys = c.y0 + np.arange(c.ny + 1) * c.dx
                    X, Y = np.meshgrid(0.5 * (xs[:-1] + xs[1:]), 0.5 * (ys[:-1] + ys[1:]))
                    rec = np.stack([X.reshape(-1), Y.reshape(-1), fld[3 * c.n_solid + 3]], 1)
                    nat.write_records(os.path.join(self.dir, f"{self.run_name}_B_{idx}.q_2d"), rec,
                                      nx, nat.FMT_E20_12, self.n_threads)
                for kind, tag, on in ((nat.SNAP_ERO2D, "erosion", c.erosion_coeff > 0.0),
                                      (nat.SNAP_ERD2D, "erodible", c.erosion_coeff > 0.0),
                                      (nat.SNAP_DEP2D, "deposit", c.settling_flag)):
                    if on:
                        rec = self.bufs[kind].a.reshape(N, -1)
                        nat.write_records(os.path.join(self.dir, f"{self.run_name}_{tag}_{idx}.q_2d"),
                                          rec, nx, nat.FMT_E20_12, self.n_threads)
            if self.phys:
                rec = self.bufs[nat.SNAP_P2D].a.reshape(N, -1)
                nat.write_records(os.path.join(self.dir, f"{self.run_name}_{idx}.p_2d"), rec, nx,
                                  nat.FMT_ES15_7, self.n_threads)
            if with_restart and self.restart:
                ns = c.n_solid
                fld = self.bufs[nat.SNAP_FIELDS].a.reshape(3 * ns + 4, N)
                B = np.ascontiguousarray(fld[3 * ns + 3])
                erodible = np.ascontiguousarray(fld[2 * ns + 3:3 * ns + 3].T)   # erodible(n_solid, nx, ny)
                nat.write_restart(os.path.join(self.dir, "restart.bin"), c.nx, c.ny, c.n_vars, ns,
                                  time, dt, self.bufs[nat.SNAP_Q].a, B, erodible, fld[:ns],
                                  fld[ns + 3:2 * ns + 3], fld[ns], fld[ns + 1], fld[ns + 2])
        except Exception as e:  # reported by the next call or by finish()
            self._error = e
        self.timings[-1][1] = _t.perf_counter() - t0

    def __call__(self, time: float, restart: bool = True):
        import threading
        import time as _t
        t0 = _t.perf_counter()
        self.finish()
        idx = lettera(self.output_idx)
        self.output_idx += 1
        what = self.what if restart else self.what & ~self.native.SNAP_Q
        self.s.snapshot_begin(what, {k: b.a for k, b in self.bufs.items()})
        self.timings.append([0.0, 0.0])
        args = (idx, time, self.s.dt, restart)
        if self.overlap:
            self._thread = threading.Thread(target=self._files, args=args, daemon=True)
            self._thread.start()
        else:
            self._files(*args)
        self.t_output = time + self.dt_output
        self.timings[-1][0] = _t.perf_counter() - t0

    def finish(self):
        if self._thread is not None:
            self._thread.join()
            self._thread = None
        if self._error is not None:
            e, self._error = self._error, None
            raise e

    def close(self):
        self.finish()
        for b in self.bufs.values():
            b.free()
        self.bufs = {}


def run_with_outputs(solver, t_end: float, runout: RunoutOutput = None, probes: ProbesOutput = None,
                     solution: SolutionOutput = None, chunk: int = 64, max_steps: int = 10 ** 8):
    """The driver loop's output schedule (src/IMEX_SfloW2D.f90:370-377,
    :395-591) over device-driven steps: the initial outputs at t_start, then
    steps up to the next output time (the dt clip of :417-433 makes the
    step that reaches it end exactly there), the outputs due, and the steady
    stop (t_steady).  Returns (steps taken, stop_flag)."""
    stop_flag = False
    t_steady = t_end
    if runout is not None:
        stop_flag = runout(solver.t, stop_flag)
    if solution is not None:
        solution(solver.t, restart=False)
    if probes is not None:
        probes(solver.t)
    steps = 0
    while solver.t < t_end and solver.t < t_steady and steps < max_steps:
        if solution is not None and t_end - solution.t_output < 1.0e-7:
            solution.t_output = t_end
        if runout is not None and t_end - runout.t_runout < 1.0e-7:
            runout.t_runout = t_end
        if probes is not None and t_end - probes.t_probes < 1.0e-7:
            probes.t_probes = t_end
        target = t_end
        if solution is not None:
            target = min(target, solution.t_output)
        if runout is not None:
            target = min(target, runout.t_runout)
        if probes is not None:
            target = min(target, probes.t_probes)
        dts = solver.run_steps(chunk, t_end=target)
        steps += int(np.count_nonzero(dts))
        if solver.t < target:
            continue
        if runout is not None and (solver.t >= runout.t_runout or solver.t >= t_steady):
            old = stop_flag
            stop_flag = runout(solver.t, stop_flag)
            if stop_flag and not old:
                t_out = solution.t_output if solution is not None else t_end
                t_steady = min(t_end, t_out)
        if probes is not None and solver.t >= probes.t_probes:
            probes(solver.t)
        if solution is not None and (solver.t >= solution.t_output or solver.t >= t_end):
            solution(solver.t)
    if solution is not None:
        solution.finish()
    return steps, stop_flag
