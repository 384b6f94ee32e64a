"""ctypes access to oracle/_ref/liboracle.so, the C restatement of the
IMEX-RK stage (oracle/imex_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  Product code never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import refpy

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib_path() -> str:
    return os.path.join(_HERE, "_ref", "liboracle.so")


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "oracle"], check=True,
                   stdout=subprocess.DEVNULL)


def _lib():
    global _LIB
    if _LIB is None:
        p = lib_path()
        if not os.path.exists(p):
            build()
        lib = ctypes.CDLL(p)
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [_I, _D]
        for name in ("oracle_destroy", "oracle_set_static", "oracle_init_state",
                     "oracle_run", "oracle_get_q", "oracle_set_q", "oracle_get_fields",
                     "oracle_get_vuln", "oracle_get_mask", "oracle_get_counters", "oracle_get_stale_cells", "oracle_set_fresh",
                     "oracle_set_time", "oracle_watch_cell",
                     "oracle_qc_to_qp", "oracle_qp_to_qc", "oracle_mixt_var",
                     "oracle_fluxes", "oracle_speeds", "oracle_implicit",
                     "oracle_implicit_cplx", "oracle_semi_impl", "oracle_expl",
                     "oracle_settling", "oracle_mass_exchange", "oracle_newton"):
            getattr(lib, name).restype = ctypes.c_int if name in (
                "oracle_run", "oracle_newton") else None
        lib.oracle_limit_minmod.restype = ctypes.c_double
        lib.oracle_pow_cr.restype = ctypes.c_double
        lib.oracle_pow_cr.argtypes = [ctypes.c_double, ctypes.c_double]
        lib.oracle_set_pow_mode.restype = None
        for name in ("oracle_mass_exchange_ers", "oracle_set_erodible", "oracle_set_noise",
                     "oracle_get_erosion",
                     "oracle_set_bcent", "oracle_get_bcent"):
            getattr(lib, name).restype = None
        lib.oracle_set_t_end.restype = None
        lib.oracle_set_t_end.argtypes = [ctypes.c_void_p, ctypes.c_double]
        lib.oracle_get_t.restype = ctypes.c_double
        lib.oracle_get_t.argtypes = [ctypes.c_void_p]
        lib.oracle_get_nan_cell.restype = ctypes.c_long
        lib.oracle_get_nan_cell.argtypes = [ctypes.c_void_p]
        lib.oracle_work_reset.restype = None
        lib.oracle_watched_faces.restype = ctypes.c_int
        lib.oracle_work_stats.restype = None
        _LIB = lib
    return _LIB


WORK_FIELDS = ("newton_calls", "newton_its", "newton_max_its", "evalf", "jac", "ls_trials",
               "cell_evals_max", "settle_calls", "settle_its", "settle_max_its",
               "hist_1", "hist_2", "hist_3", "hist_4", "hist_5_8", "hist_9_16", "hist_17_32",
               "hist_gt32")


def work_reset() -> None:
    """Zero the process-wide Newton / settling work counters."""
    _lib().oracle_work_reset()


def work_stats() -> dict:
    """The work counters since work_reset() (see oracle_work_stats)."""
    out = np.zeros(len(WORK_FIELDS), np.int64)
    _lib().oracle_work_stats(out.ctypes.data_as(ctypes.POINTER(ctypes.c_long)))
    return dict(zip(WORK_FIELDS, (int(v) for v in out)))


def pow_cr(x: float, y: float) -> float:
    """x**y correctly rounded (quad-precision pow rounded once)."""
    return _lib().oracle_pow_cr(x, y)


def limit_minmod(v3, z3) -> float:
    v = np.ascontiguousarray(v3, np.float64)
    z = np.ascontiguousarray(z3, np.float64)
    return _lib().oracle_limit_minmod(_dp(v), _dp(z))


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class Oracle:
    """One oracle run (the C restatement).  Arrays use [k, j, var] order
    (identical memory to the reference's q(var, j, k))."""

    def __init__(self, cfg, static=None, q0=None):
        self.lib = _lib()
        self.cfg = cfg
        ip, rp = refpy.pack_params(cfg)
        self.ip, self.rp = ip, rp
        self.h = ctypes.c_void_p(self.lib.oracle_create(_ip(ip), _dp(rp)))
        self.nv = cfg.n_vars
        if static is not None:
            self.set_static(static)
            if static.get("erodible") is not None:
                self.set_erodible(static["erodible"])
            if static.get("Z") is not None and getattr(cfg, "stochastic_flag", False):
                self.set_noise(static["Z"])
            if static.get("B_cent") is not None:
                b = np.ascontiguousarray(static["B_cent"], np.float64)
                self.lib.oracle_set_bcent(self.h, _dp(b))
        if q0 is not None:
            self.init_state(q0)

    def __del__(self):
        try:
            if self.h:
                self.lib.oracle_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_static(self, st):
        c = self.cfg
        N = c.nx * c.ny

        def a(name, dt, default=0):
            v = st.get(name)
            if v is None:
                return np.full(N, default, dtype=dt)
            return np.ascontiguousarray(v, dtype=dt).reshape(-1)

        self._st = [a("B_prime_x", np.float64), a("B_prime_y", np.float64),
                    a("source_cell", np.int32), a("source_sides", np.int32),
                    a("arc_perim", np.float64), a("arc_n_x", np.float64),
                    a("arc_n_y", np.float64)]
        bx, by, sc, sd, ap, ax, ay = self._st
        self.lib.oracle_set_static(self.h, _dp(bx), _dp(by), _ip(sc), _ip(sd), _dp(ap),
                                   _dp(ax), _dp(ay), ctypes.c_double(st.get("h_source", 0.0)))

    def init_state(self, q):
        a = np.ascontiguousarray(q, dtype=np.float64)
        assert a.shape == (self.cfg.ny, self.cfg.nx, self.nv)
        self.lib.oracle_init_state(self.h, _dp(a))

    def run(self, n):
        dts = np.zeros(max(n, 1))
        nc = np.zeros(max(n, 1), np.int32)
        err = self.lib.oracle_run(self.h, ctypes.c_int(n), _dp(dts), _ip(nc))
        return dts[:n], nc[:n], err

    def q(self):
        out = np.zeros((self.cfg.ny, self.cfg.nx, self.nv))
        self.lib.oracle_get_q(self.h, _dp(out))
        return out

    def set_q(self, q):
        a = np.ascontiguousarray(q, dtype=np.float64)
        self.lib.oracle_set_q(self.h, _dp(a))

    def fields(self):
        c = self.cfg
        hm = np.zeros((c.ny, c.nx))
        pd = np.zeros((c.ny, c.nx))
        mv = np.zeros((c.ny, c.nx))
        dep = np.zeros((c.n_solid, c.ny, c.nx))
        t = ctypes.c_double()
        dt = ctypes.c_double()
        self.lib.oracle_get_fields(self.h, _dp(hm), _dp(pd), _dp(mv), _dp(dep),
                                   ctypes.byref(t), ctypes.byref(dt))
        return dict(hmax=hm, pdynmax=pd, mod_vel_max=mv, deposit=dep, t=t.value, dt=dt.value)

    def vuln(self):
        c = self.cfg
        nl = max(1, len(c.thickness_levels) * len(c.dyn_pres_levels))
        v = np.zeros((c.ny, c.nx, nl), np.int32)
        self.lib.oracle_get_vuln(self.h, _ip(v))
        return v

    def mask(self):
        m = np.zeros((self.cfg.ny, self.cfg.nx), np.int32)
        self.lib.oracle_get_mask(self.h, _ip(m))
        return m

    def set_fresh(self, fresh: bool):
        """True: states the reference keeps from an earlier step (dry cells
        that left the active set) are read as a fresh evaluation gives them,
        as on the device."""
        self.lib.oracle_set_fresh(self.h, 1 if fresh else 0)

    def set_pow_mode(self, mode: int):
        """0: the C library's pow (the reference's); 1: correctly rounded."""
        self.lib.oracle_set_pow_mode(self.h, ctypes.c_int(mode))

    def set_time(self, t, dt_old, dt_old_old):
        self.lib.oracle_set_time(self.h, ctypes.c_double(t), ctypes.c_double(dt_old),
                                 ctypes.c_double(dt_old_old))

    def mass_exchange_ers(self, qp, erod, dt):
        """(deposition, eqns_term, erosion) rates with the cells' erodible
        layers erod [n, n_solid]."""
        qp = np.ascontiguousarray(qp, np.float64)
        erod = np.ascontiguousarray(erod, np.float64)
        n = qp.shape[0]
        dep = np.zeros((n, self.cfg.n_solid))
        eq = np.zeros((n, self.nv))
        ers = np.zeros((n, self.cfg.n_solid))
        self.lib.oracle_mass_exchange_ers(self.h, ctypes.c_int(n), _dp(qp), _dp(erod),
                                          ctypes.c_double(dt), _dp(dep), _dp(eq), _dp(ers))
        return dep, eq, ers

    def set_erodible(self, e):
        """erodible as [class, k, j]."""
        a = np.ascontiguousarray(e, np.float64)
        assert a.shape == (self.cfg.n_solid, self.cfg.ny, self.cfg.nx)
        self.lib.oracle_set_erodible(self.h, _dp(a))

    def set_noise(self, Z):
        """The cells' stochastic values Z as [k, j]."""
        a = np.ascontiguousarray(Z, np.float64)
        assert a.shape == (self.cfg.ny, self.cfg.nx)
        self.lib.oracle_set_noise(self.h, _dp(a))

    def erosion(self):
        """(erosion, erodible) as [class, k, j]."""
        c = self.cfg
        ers = np.zeros((c.n_solid, c.ny, c.nx))
        erd = np.zeros((c.n_solid, c.ny, c.nx))
        self.lib.oracle_get_erosion(self.h, _dp(ers), _dp(erd))
        return ers, erd

    def topography(self):
        """B_cent as [k, j]."""
        b = np.zeros((self.cfg.ny, self.cfg.nx))
        self.lib.oracle_get_bcent(self.h, _dp(b))
        return b

    def set_t_end(self, t_end):
        self.lib.oracle_set_t_end(self.h, ctypes.c_double(t_end))

    @property
    def t(self):
        return self.lib.oracle_get_t(self.h)

    def nan_cell(self):
        """(j, k) of the first cell whose assembled state held a NaN in the
        last run() call (the reference prints it and pauses), or None."""
        c = self.lib.oracle_get_nan_cell(self.h)
        return None if c < 0 else (int(c % self.cfg.nx), int(c // self.cfg.nx))

    def watch_cell(self, j, k):
        """Keep the KT fluxes of the west and east x faces of cell (j, k)."""
        self.lib.oracle_watch_cell(self.h, ctypes.c_int(j), ctypes.c_int(k))

    def watched_faces(self):
        """(pre-cap fluxes [2, n_vars], post fluxes [2, n_vars], faces seen bits)."""
        pre = np.zeros(16)
        post = np.zeros(16)
        seen = self.lib.oracle_watched_faces(self.h, _dp(pre), _dp(post))
        nv = self.nv
        return pre.reshape(2, 8)[:, :nv], post.reshape(2, 8)[:, :nv], seen

    def stale_cells(self):
        """[k, j] bool: cells next to a stale read that differed from a fresh
        evaluation since creation."""
        nx, ny = self.cfg.nx, self.cfg.ny
        out = np.zeros(nx * ny, np.uint8)
        self.lib.oracle_get_stale_cells(self.h, out.ctypes.data_as(ctypes.c_void_p))
        return out.reshape(ny, nx).astype(bool)

    def counters(self):
        d = ctypes.c_long()
        s = ctypes.c_long()
        self.lib.oracle_get_counters(self.h, ctypes.byref(d), ctypes.byref(s))
        return d.value, s.value

    # ---- per-function ------------------------------------------------------
    def qc_to_qp(self, qc):
        qc = np.ascontiguousarray(qc, np.float64)
        n = qc.shape[0]
        qp = np.zeros((n, self.nv + 2))
        pd = np.zeros(n)
        self.lib.oracle_qc_to_qp(self.h, ctypes.c_int(n), _dp(qc), _dp(qp), _dp(pd))
        return qp, pd

    def qp_to_qc(self, qp):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        qc = np.zeros((n, self.nv))
        self.lib.oracle_qp_to_qc(self.h, ctypes.c_int(n), _dp(qp), _dp(qc))
        return qc

    def mixt_var(self, qp):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, 6))
        self.lib.oracle_mixt_var(self.h, ctypes.c_int(n), _dp(qp), _dp(out))
        return out

    def fluxes(self, qc, qp, direction):
        qc = np.ascontiguousarray(qc, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.oracle_fluxes(self.h, ctypes.c_int(n), ctypes.c_int(direction), _dp(qc),
                               _dp(qp), _dp(out))
        return out

    def speeds(self, qp, direction):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, 2))
        self.lib.oracle_speeds(self.h, ctypes.c_int(n), ctypes.c_int(direction), _dp(qp),
                               _dp(out))
        return out

    def implicit(self, qc):
        qc = np.ascontiguousarray(qc, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.oracle_implicit(self.h, ctypes.c_int(n), _dp(qc), _dp(out))
        return out

    def implicit_cplx(self, qre, qim):
        qre = np.ascontiguousarray(qre, np.float64)
        qim = np.ascontiguousarray(qim, np.float64)
        n = qre.shape[0]
        ore = np.zeros((n, self.nv))
        oim = np.zeros((n, self.nv))
        self.lib.oracle_implicit_cplx(self.h, ctypes.c_int(n), _dp(qre), _dp(qim), _dp(ore),
                                      _dp(oim))
        return ore, oim

    def semi_impl(self, qc, qp):
        qc = np.ascontiguousarray(qc, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.oracle_semi_impl(self.h, ctypes.c_int(n), _dp(qc), _dp(qp), _dp(out))
        return out

    def expl(self, bx, by, t, qp):
        bx = np.ascontiguousarray(bx, np.float64)
        by = np.ascontiguousarray(by, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.oracle_expl(self.h, ctypes.c_int(n), _dp(bx), _dp(by), ctypes.c_double(t),
                             _dp(qp), _dp(out))
        return out

    def settling(self, diam, rhos, rhoc, ikv):
        a = [np.ascontiguousarray(x, np.float64) for x in (diam, rhos, rhoc, ikv)]
        n = a[0].shape[0]
        out = np.zeros(n)
        self.lib.oracle_settling(self.h, ctypes.c_int(n), *[_dp(x) for x in a], _dp(out))
        return out

    def mass_exchange(self, qp, dt):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        dep = np.zeros((n, self.cfg.n_solid))
        eq = np.zeros((n, self.nv))
        self.lib.oracle_mass_exchange(self.h, ctypes.c_int(n), _dp(qp), ctypes.c_double(dt),
                                      _dp(dep), _dp(eq))
        return dep, eq

    def newton(self, dt, q_si, q0, R):
        a = [np.ascontiguousarray(x, np.float64) for x in (q_si, q0, R)]
        out = np.zeros(self.nv)
        it = self.lib.oracle_newton(self.h, ctypes.c_double(dt), _dp(a[0]), _dp(a[1]),
                                    _dp(a[2]), _dp(out))
        return out, it
