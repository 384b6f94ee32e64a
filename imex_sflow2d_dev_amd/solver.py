"""Host-side mirror of the reference's MODULE solver_2d (src/solver_2d.f90)
over the device-resident HIP implementation.

Method names and their meaning follow the reference's public procedures
(check_solve, timestep, imex_RK_solver, update_erosion_deposition_cell) and
the driver loop of PROGRAM IMEX_SfloW2D (src/IMEX_SfloW2D.f90:395-505).  The
state never leaves the GPU between steps; the q / hmax / ... properties copy
it back on demand (the reference's module arrays are read the same way by its
output routines).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import native
from .config import RunConfig


class PhysicsError(RuntimeError):
    """A condition on which the reference prints diagnostics and pauses."""

    def __init__(self, bits: int):
        names = [n for n, b in (("NaN", 1), ("negative thickness", 2),
                                ("negative solid mass", 4), ("T < 273 K", 8),
                                ("negative deposition", 16),
                                ("solid mass > total mass", 32)) if bits & b]
        super().__init__("physics error: " + ", ".join(names))
        self.bits = bits


class Solver:
    def __init__(self, cfg: RunConfig, static: dict, q0: np.ndarray, t: float = None,
                 raise_on_error: bool = True):
        self.cfg = cfg
        self.nv = cfg.n_vars
        if static.get("h_source") is not None:
            # the source thickness the set-up derived from the mass flow rate
            # (src/inpout_2d.f90:4522-4531) is part of the configuration
            cfg.h_source = float(static["h_source"])
        self.H = native.Handle(cfg)
        self.L = self.H.L
        self.h = self.H.h
        self.raise_on_error = raise_on_error
        self.t = cfg.t_start if t is None else t
        self.dt = cfg.dt0
        self.dt_old = cfg.dt0
        self.dt_old_old = cfg.dt0
        self.err = 0
        self._host_driven = False  # host-driven steps since the device last had t
        self._set_static(static)
        self._upload(q0)

    # ---- set-up -------------------------------------------------------------
    def _set_static(self, st: dict):
        N = self.cfg.nx * self.cfg.ny

        def d(name):
            v = st.get(name)
            if v is None:
                return None
            return np.ascontiguousarray(v, dtype=np.float64).reshape(N)

        def i(name):
            v = st.get(name)
            if v is None:
                return None
            return np.ascontiguousarray(v, dtype=np.int32).reshape(N)

        self._static = [d("B_prime_x"), d("B_prime_y"), i("source_cell"), i("source_sides"),
                        d("arc_perim"), d("arc_n_x"), d("arc_n_y")]
        s = native.ImexStatic()
        names = ["B_prime_x", "B_prime_y", "source_cell", "source_sides", "arc_perim",
                 "arc_n_x", "arc_n_y"]
        for name, arr in zip(names, self._static):
            if arr is not None:
                ptr = native._ip(arr) if arr.dtype == np.int32 else native._dp(arr)
                setattr(s, name, ptr)
        self.H.check(self.L.imex_set_static(self.h, ctypes.byref(s)), "imex_set_static")
        if st.get("erodible") is not None and (self.cfg.erosion_coeff > 0.0
                                               or self.cfg.erodible_deposit_flag):
            self.set_erodible(st["erodible"])
        if self.cfg.stochastic_flag and st.get("Z") is not None:
            self.set_noise(st["Z"])
        if self.cfg.topo_change_flag:
            if st.get("B_cent") is None:
                raise ValueError("topo_change_flag needs B_cent in the static fields")
            self.set_topography(st["B_cent"])

    def set_erodible(self, e: np.ndarray):
        """The erodible layer as [class, k, j]."""
        a = np.ascontiguousarray(np.transpose(np.asarray(e, dtype=np.float64), (1, 2, 0)))
        assert a.shape == (self.cfg.ny, self.cfg.nx, self.cfg.n_solid)
        self.H.check(self.L.imex_set_erodible(self.h, native._dp(a)), "imex_set_erodible")

    def set_noise(self, Z: np.ndarray):
        """The cells' stochastic values Z as [k, j] (stochastic_flag)."""
        a = np.ascontiguousarray(Z, dtype=np.float64)
        assert a.shape == (self.cfg.ny, self.cfg.nx)
        self.H.check(self.L.imex_set_noise(self.h, native._dp(a)), "imex_set_noise")

    def topography(self) -> np.ndarray:
        """B_cent as [k, j] (evolves with topo_change_flag)."""
        b = np.zeros((self.cfg.ny, self.cfg.nx))
        self.H.check(self.L.imex_get_topography(self.h, native._dp(b)), "imex_get_topography")
        return b

    def erosion(self):
        """(erosion, erodible) as [class, k, j]."""
        c = self.cfg
        ers = np.zeros((c.n_solid, c.ny, c.nx))
        erd = np.zeros((c.ny, c.nx, c.n_solid))
        self.H.check(self.L.imex_get_erosion(self.h, native._dp(ers), native._dp(erd)),
                     "imex_get_erosion")
        return ers, np.transpose(erd, (2, 0, 1)).copy()

    def _upload(self, q0: np.ndarray):
        a = np.ascontiguousarray(q0, dtype=np.float64)
        assert a.shape == (self.cfg.ny, self.cfg.nx, self.nv), a.shape
        self.H.check(self.L.imex_upload_state(self.h, native._dp(a), ctypes.c_double(self.t),
                                              ctypes.c_double(self.cfg.dt0)), "imex_upload_state")

    # ---- reference procedures ------------------------------------------------
    def check_solve(self, solve_all: bool = False):
        self.H.check(self.L.imex_check_solve(self.h, ctypes.c_int(int(solve_all))), "check_solve")

    def timestep(self) -> float:
        dt = ctypes.c_double()
        self.H.check(self.L.imex_timestep(self.h, ctypes.c_double(self.t), ctypes.byref(dt)),
                     "timestep")
        self.dt = dt.value
        return self.dt

    def imex_RK_solver(self):
        self.H.check(self.L.imex_rk_step(self.h, ctypes.c_double(self.dt)), "imex_RK_solver")

    def update_erosion_deposition_cell(self, dt: float):
        self.H.check(self.L.imex_erosion_deposition(self.h, ctypes.c_double(dt)),
                     "update_erosion_deposition_cell")

    def post_step(self, want_stats: bool = False):
        st = native.ImexStepStats()
        rc = self.L.imex_post_step(self.h, ctypes.c_double(self.t),
                                   ctypes.byref(st) if want_stats else None)
        self.H.check(rc, "post_step")
        self._raise(rc)
        return st if want_stats else None

    def _raise(self, rc: int):
        if rc > 0:
            self.err |= rc
            if self.raise_on_error:
                raise PhysicsError(rc)

    # ---- driver --------------------------------------------------------------
    def step(self, t_end: float = None):
        """One pass of the driver loop body with the host-side dt clip."""
        t_end = self.cfg.t_end if t_end is None else t_end
        self.check_solve(False)
        dt = self.timestep()
        if self.t + dt > t_end:
            dt = t_end - self.t
        dt = min(dt, 1.1 * 0.5 * (self.dt_old + self.dt_old_old))
        self.dt_old_old = self.dt_old
        self.dt_old = dt
        self._host_driven = True
        self.dt = dt
        self.imex_RK_solver()
        self.update_erosion_deposition_cell(dt)
        if self.cfg.topo_change_flag:
            self.H.check(self.L.imex_topography_reconstruction(self.h), "topography_reconstruction")
        self.t = self.t + dt
        self.post_step()
        return dt

    def run_steps(self, nsteps: int, t_end: float = None, record_dt: bool = True):
        """nsteps driver steps with the whole loop on the device."""
        t_end = self.cfg.t_end if t_end is None else t_end
        self._sync_time_to_device()
        dts = np.zeros(max(nsteps, 1)) if record_dt else None
        rc = self.L.imex_run_steps(self.h, ctypes.c_int(nsteps), ctypes.c_double(t_end),
                                   native._dp(dts) if record_dt else None)
        self.H.check(rc, "imex_run_steps")
        self._pull_time()
        self._raise(rc)
        return dts[:nsteps] if record_dt else None

    def run_steps_timed(self, nsteps: int, t_end: float = None, record_dt: bool = True,
                        sample_every: int = 1):
        """run_steps with HIP events around the launches of every
        sample_every-th step: returns (dts or None, kernel_ms[3] summed over
        the sampled steps, number of sampled steps, active cells summed over
        all steps) -- the three kernels are the active set, the fluxes and
        the update (stage 2, sedimentation, epilogue)."""
        t_end = self.cfg.t_end if t_end is None else t_end
        self._sync_time_to_device()
        dts = np.zeros(max(nsteps, 1)) if record_dt else None
        kms = np.zeros(3)
        ns = ctypes.c_int()
        cells = np.zeros(1, np.int64)
        rc = self.L.imex_run_steps_timed(self.h, ctypes.c_int(nsteps), ctypes.c_double(t_end),
                                         native._dp(dts) if record_dt else None,
                                         ctypes.c_int(sample_every), native._dp(kms),
                                         ctypes.byref(ns),
                                         cells.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        self.H.check(rc, "imex_run_steps_timed")
        self._pull_time()
        self._raise(rc)
        return (dts[:nsteps] if record_dt else None), kms, ns.value, int(cells[0])

    def _sync_time_to_device(self):
        # the device-side driver keeps t and the dt history in device memory;
        # host-driven steps advance them on the host, so push them first
        if self._host_driven:
            self.H.check(self.L.imex_set_time(self.h, ctypes.c_double(self.t),
                                              ctypes.c_double(self.dt_old),
                                              ctypes.c_double(self.dt_old_old)), "imex_set_time")
            self._host_driven = False

    def _pull_time(self):
        t, d1, d2, dt = (ctypes.c_double() for _ in range(4))
        self.H.check(self.L.imex_get_time_history(self.h, ctypes.byref(t), ctypes.byref(d1),
                                                  ctypes.byref(d2)), "imex_get_time_history")
        self.H.check(self.L.imex_get_time(self.h, None, ctypes.byref(dt)), "imex_get_time")
        self.t, self.dt_old, self.dt_old_old, self.dt = t.value, d1.value, d2.value, dt.value

    def _download(self, what, **kw):
        c = self.cfg
        bufs = dict(q=None, hmax=None, pdynmax=None, mvel=None, dep=None, vuln=None, mask=None)
        bufs.update(kw)
        ptr = lambda a, f: f(a) if a is not None else None
        self.H.check(self.L.imex_download(
            self.h, ctypes.c_uint(what), ptr(bufs["q"], native._dp), ptr(bufs["hmax"], native._dp),
            ptr(bufs["pdynmax"], native._dp), ptr(bufs["mvel"], native._dp),
            ptr(bufs["dep"], native._dp), ptr(bufs["vuln"], native._ip),
            ptr(bufs["mask"], native._ip)), "imex_download")

    @property
    def q(self) -> np.ndarray:
        out = np.zeros((self.cfg.ny, self.cfg.nx, self.nv))
        self._download(native.DL_Q, q=out)
        return out

    def fields(self) -> dict:
        c = self.cfg
        hm = np.zeros((c.ny, c.nx))
        pd = np.zeros((c.ny, c.nx))
        mv = np.zeros((c.ny, c.nx))
        dep = np.zeros((c.n_solid, c.ny, c.nx))
        self._download(native.DL_STATS | native.DL_DEPOSIT, hmax=hm, pdynmax=pd, mvel=mv, dep=dep)
        return dict(hmax=hm, pdynmax=pd, mod_vel_max=mv, deposit=dep, t=self.t, dt=self.dt)

    def vuln(self) -> np.ndarray:
        c = self.cfg
        nl = max(1, len(c.thickness_levels) * len(c.dyn_pres_levels))
        v = np.zeros((c.ny, c.nx, nl), np.int32)
        if len(c.thickness_levels) * len(c.dyn_pres_levels) > 0:
            self._download(native.DL_VULN, vuln=v)
        return v

    def mask(self) -> np.ndarray:
        m = np.zeros((self.cfg.ny, self.cfg.nx), np.int32)
        self._download(native.DL_MASK, mask=m)
        return m

    def counters(self):
        """(active cells summed over every check_solve since create,
        device-driven steps since the upload)."""
        a = ctypes.c_int64()
        n = ctypes.c_int64()
        self.H.check(self.L.imex_get_counters(self.h, ctypes.byref(a), ctypes.byref(n)),
                     "imex_get_counters")
        return a.value, n.value

    # ---- device-side outputs (the grid work of output_runout / output_probes)
    def set_topography(self, B: np.ndarray):
        """B_cent as [k, j]; output_runout reads it when the runout origin is
        not set (src/inpout_2d.f90:7293)."""
        b = np.ascontiguousarray(B, dtype=np.float64)
        assert b.shape == (self.cfg.ny, self.cfg.nx)
        self.H.check(self.L.imex_set_topography(self.h, native._dp(b)), "imex_set_topography")

    def runout_reduce(self, x0: float, y0: float, alphas_threshold: float = 0.0,
                      store: bool = False, with_b: bool = False) -> dict:
        """The whole-array expressions of one output_runout call on the device
        state (imex_runout_reduce); store=True then sets hpos_old := hpos."""
        r = native.ImexRunoutRed()
        flags = (native.RUNOUT_STORE if store else 0) | (native.RUNOUT_B if with_b else 0)
        self.H.check(self.L.imex_runout_reduce(
            self.h, ctypes.c_double(x0), ctypes.c_double(y0), ctypes.c_double(alphas_threshold),
            ctypes.c_uint(flags), ctypes.byref(r)), "imex_runout_reduce")
        return r.as_dict()

    def probes(self, xy) -> np.ndarray:
        """output_probes' values per probe ([n, 6 + n_solid (+ n_solid with
        settling)]) at the points xy ([n, 2])."""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        c = self.cfg
        out = np.zeros((len(xy), 6 + c.n_solid + (c.n_solid if c.settling_flag else 0)))
        self.H.check(self.L.imex_probes(self.h, ctypes.c_int(len(xy)), native._dp(xy),
                                        native._dp(out)), "imex_probes")
        return out

    def snapshot_cols(self, kind: int) -> int:
        return self.H.check(self.L.imex_snapshot_cols(self.h, ctypes.c_uint(kind)),
                            "imex_snapshot_cols")

    def snapshot_begin(self, what: int, bufs: dict):
        """Starts a snapshot (imex_snapshot_begin) into the host arrays of
        `bufs` (keys native.SNAP_*: float64 arrays of N * cols values,
        page-locked for the copy to overlap the next steps)."""
        N = self.cfg.nx * self.cfg.ny
        ptrs = []
        for kind in native.SNAP_KINDS:
            a = bufs.get(kind) if what & kind else None
            if a is not None:
                assert a.dtype == np.float64 and a.size >= N * self.snapshot_cols(kind)
                ptrs.append(native._dp(a))
            else:
                ptrs.append(None)
        self.H.check(self.L.imex_snapshot_begin(self.h, ctypes.c_uint(what), *ptrs),
                     "imex_snapshot_begin")

    def snapshot_wait(self):
        self.H.check(self.L.imex_snapshot_wait(self.h), "imex_snapshot_wait")

    def footprint(self) -> int:
        """Bytes of device memory the run holds."""
        b = ctypes.c_size_t()
        self.H.check(self.L.imex_get_footprint(self.h, ctypes.byref(b)), "imex_get_footprint")
        return int(b.value)

    def work_counters(self):
        """(cells the update kernel processed summed over the steps since
        create, cells of the always-active boundary frame per check_solve)."""
        w, f = ctypes.c_int64(), ctypes.c_int64()
        self.H.check(self.L.imex_get_work_counters(self.h, ctypes.byref(w), ctypes.byref(f)),
                     "imex_get_work_counters")
        return w.value, f.value

    def stats(self):
        st = native.ImexStepStats()
        self.H.check(self.L.imex_post_step(self.h, ctypes.c_double(self.t), ctypes.byref(st)),
                     "stats")
        return st

    def close(self):
        self.H.close()


class KernelProbe:
    """Per-function device evaluations (for known-answer tests)."""

    def __init__(self, cfg: RunConfig):
        self.cfg = cfg
        self.nv = cfg.n_vars
        self.H = native.Handle(cfg)
        self.L = self.H.L

    def _run(self, name, *args):
        self.H.check(getattr(self.L, name)(self.H.h, *args), name)

    def qc_to_qp(self, qc):
        qc = np.ascontiguousarray(qc, np.float64)
        n = qc.shape[0]
        qp = np.zeros((n, self.nv + 2))
        pd = np.zeros(n)
        self._run("imex_test_qc_to_qp", ctypes.c_int(n), native._dp(qc), native._dp(qp),
                  native._dp(pd))
        return qp, pd

    def qp_to_qc(self, qp):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        qc = np.zeros((n, self.nv))
        self._run("imex_test_qp_to_qc", ctypes.c_int(n), native._dp(qp), native._dp(qc))
        return qc

    def fluxes(self, qc, qp, direction):
        qc = np.ascontiguousarray(qc, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self._run("imex_test_fluxes", ctypes.c_int(n), ctypes.c_int(direction), native._dp(qc),
                  native._dp(qp), native._dp(out))
        return out

    def speeds(self, qp, direction):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, 2))
        self._run("imex_test_speeds", ctypes.c_int(n), ctypes.c_int(direction), native._dp(qp),
                  native._dp(out))
        return out

    def implicit_cplx(self, qre, qim):
        qre = np.ascontiguousarray(qre, np.float64)
        qim = np.ascontiguousarray(qim, np.float64)
        n = qre.shape[0]
        ore = np.zeros((n, self.nv))
        oim = np.zeros((n, self.nv))
        self._run("imex_test_implicit_cplx", ctypes.c_int(n), native._dp(qre), native._dp(qim),
                  native._dp(ore), native._dp(oim))
        return ore, oim

    def implicit_real(self, q):
        """Implicit terms at real states as the Newton residual evaluates
        them; returns (terms, fast) with fast[i] True where the real-arithmetic
        evaluation was used."""
        q = np.ascontiguousarray(q, np.float64)
        n = q.shape[0]
        out = np.zeros((n, self.nv))
        fast = np.zeros(n, np.int32)
        self._run("imex_test_implicit_real", ctypes.c_int(n), native._dp(q), native._dp(out),
                  native._ip(fast))
        return out, fast.astype(bool)

    def semi_impl(self, qp):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, self.nv))
        self._run("imex_test_semi_impl", ctypes.c_int(n), native._dp(qp), native._dp(out))
        return out

    def expl(self, bx, by, t, qp):
        bx = np.ascontiguousarray(bx, np.float64)
        by = np.ascontiguousarray(by, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, self.nv))
        self._run("imex_test_expl", ctypes.c_int(n), native._dp(bx), native._dp(by),
                  ctypes.c_double(t), native._dp(qp), native._dp(out))
        return out

    def settling(self, diam, rhos, rhoc, ikv):
        a = [np.ascontiguousarray(x, np.float64) for x in (diam, rhos, rhoc, ikv)]
        n = a[0].shape[0]
        out = np.zeros(n)
        self._run("imex_test_settling", ctypes.c_int(n), *[native._dp(x) for x in a],
                  native._dp(out))
        return out

    def mass_exchange(self, qp, dt):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        dep = np.zeros((n, self.cfg.n_solid))
        eq = np.zeros((n, self.nv))
        self._run("imex_test_mass_exchange", ctypes.c_int(n), native._dp(qp), ctypes.c_double(dt),
                  native._dp(dep), native._dp(eq))
        return dep, eq

    def mass_exchange_ers(self, qp, erod, dt):
        """(deposition, eqns_term, erosion) rates with the cells' erodible
        layers erod [n, n_solid]."""
        qp = np.ascontiguousarray(qp, np.float64)
        erod = np.ascontiguousarray(erod, np.float64)
        n = qp.shape[0]
        dep = np.zeros((n, self.cfg.n_solid))
        eq = np.zeros((n, self.nv))
        ers = np.zeros((n, self.cfg.n_solid))
        self._run("imex_test_mass_exchange_ers", ctypes.c_int(n), native._dp(qp), native._dp(erod),
                  ctypes.c_double(dt), native._dp(dep), native._dp(eq), native._dp(ers))
        return dep, eq, ers

    def newton(self, dt, q_si, q0, R):
        a = [np.ascontiguousarray(x, np.float64) for x in (q_si, q0, R)]
        n = a[0].shape[0]
        out = np.zeros((n, self.nv))
        its = np.zeros(n, np.int32)
        self._run("imex_test_newton", ctypes.c_int(n), ctypes.c_double(dt),
                  *[native._dp(x) for x in a], native._dp(out), native._ip(its))
        return out, its

    def pow(self, x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(y, np.float64)
        out = np.zeros_like(x)
        self._run("imex_test_pow", ctypes.c_int(len(x)), native._dp(x), native._dp(y),
                  native._dp(out))
        return out

    def arith(self, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        n = a.shape[0]
        out = np.zeros((n, 4))
        self._run("imex_test_arith", ctypes.c_int(n), native._dp(a), native._dp(b), native._dp(out))
        return out

    def close(self):
        self.H.close()
