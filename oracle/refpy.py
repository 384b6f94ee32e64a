"""ctypes access to oracle/_ref/libimexref.so, the reference Fortran solver
modules built from their own sources by oracle/Makefile.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by bench.py's cpu_baseline
leg, as the checker.  Product code never imports this module.

The library keeps its state in Fortran module variables, so one process holds
one reference run at a time (setup() tears the previous one down).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib_path(flavour: str = "ieee") -> str:
    name = "libimexref.so" if flavour == "ieee" else "libimexref_perf.so"
    return os.path.join(_HERE, "_ref", name)


def available(flavour: str = "ieee") -> bool:
    return os.path.exists(lib_path(flavour))


def _lib(flavour: str = "ieee"):
    if flavour not in _LIBS:
        p = lib_path(flavour)
        if not os.path.exists(p):
            raise FileNotFoundError(f"{p} not built (make -C oracle ref)")
        L = ctypes.CDLL(p)
        L.imexref_eps_sing.restype = ctypes.c_double
        _LIBS[flavour] = L
    return _LIBS[flavour]


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


def set_threads(n: int, flavour: str = "ieee") -> None:
    """OpenMP threads of the reference's cell loops (one for the pore-
    pressure parity runs: several threads race on a module variable)."""
    _lib(flavour).imexref_set_threads(ctypes.c_int(int(n)))


def pack_params(cfg):
    """ip[55] / rp[111] parameter vectors read by imexref_setup (and by the
    C restatement's oracle_create)."""
    ip = np.zeros(55, dtype=np.int32)
    ip[0] = cfg.nx
    ip[1] = cfg.ny
    ip[2] = cfg.n_solid
    ip[3] = int(cfg.energy_flag)
    ip[4] = int(cfg.settling_flag)
    ip[5] = int(cfg.entrainment_flag)
    ip[6] = int(cfg.radial_source_flag)
    ip[7] = int(cfg.collapsing_volume_flag)
    ip[8] = len(cfg.thickness_levels)
    ip[9] = len(cfg.dyn_pres_levels)
    ip[10] = cfg.limiter
    ip[11] = cfg.bc_flag
    ip[12] = int(getattr(cfg, "erodible_deposit_flag", False))
    ip[13] = int(getattr(cfg, "topo_change_flag", False))
    ip[14] = getattr(cfg, "rheology_model", 1)
    ip[15] = int(getattr(cfg, "stochastic_flag", False))
    rp = np.zeros(111, dtype=np.float64)
    rp[0:16] = [cfg.dx, cfg.x0, cfg.y0, cfg.cfl, cfg.dt0, cfg.max_dt, cfg.t_start,
                cfg.t_end, cfg.grav, cfg.mu, cfg.xi, cfg.pres, cfg.T_ambient,
                cfg.sp_heat_a, cfg.sp_gas_const_a, cfg.kin_visc_a]
    rp[16:21] = [cfg.x_collapse, cfg.y_collapse, cfg.r_collapse, cfg.h_collapse,
                 cfg.T_collapse]
    rp[21:27] = [cfg.x_source, cfg.y_source, cfg.r_source, cfg.vel_source,
                 cfg.T_source, cfg.h_source]
    rp[27:31] = cfg.time_param[:4]
    for i in range(cfg.n_solid):
        rp[31 + i] = cfg.rho_s[i]
        rp[34 + i] = cfg.diam_s[i]
        rp[37 + i] = cfg.sp_heat_s[i]
        rp[40 + i] = cfg.alphas_collapse[i]
        rp[43 + i] = cfg.alphas_source[i]
    for i, v in enumerate(cfg.thickness_levels):
        rp[46 + i] = v
    for i, v in enumerate(cfg.dyn_pres_levels):
        rp[54 + i] = v
    rp[62] = cfg.mfr_source
    rp[63] = getattr(cfg, "erosion_coeff", 0.0)
    rp[64] = getattr(cfg, "erodible_porosity", 0.0)
    rp[65] = getattr(cfg, "T_erodible", 300.0)
    for i, v in enumerate(cfg.erodible_fractions()):
        rp[66 + i] = v
    rp[69:73] = [getattr(cfg, "mu_s", 0.0), getattr(cfg, "mu_2", 0.0),
                 getattr(cfg, "muI_inf", 0.0), getattr(cfg, "I_0", 0.0)]
    # pore pressure
    ip[16] = int(getattr(cfg, "pore_pressure_flag", False))
    ip[17] = int(getattr(cfg, "gas_loss_flag", False))
    ip[18] = cfg.f_inhibit_code if hasattr(cfg, "f_inhibit_code") else 0
    ip[19] = getattr(cfg, "N_inh", 2)
    ip[20] = int(getattr(cfg, "dynamic_permeability_flag", False))
    rp[73] = getattr(cfg, "hydraulic_permeability", 0.0)
    rp[74] = getattr(cfg, "alpha_trans", 0.0)
    rp[75] = getattr(cfg, "maximum_solid_packing", 0.64)
    sph = list(getattr(cfg, "sphericity_s", [1.0, 1.0, 1.0]))
    for i in range(min(3, cfg.n_solid)):
        rp[76 + i] = sph[i] if i < len(sph) else 1.0
    # per-variable boundary conditions, sides W, E, S, N
    bcf, bcv = getattr(cfg, "bc_flags", None), getattr(cfg, "bc_values", None)
    if bcf is not None:
        ip[21] = 1
        for sd in range(4):
            for i in range(cfg.n_vars):
                ip[22 + 8 * sd + i] = bcf[sd][i]
                rp[79 + 8 * sd + i] = bcv[sd][i]
    # the transported noise
    ip[54] = int(getattr(cfg, "stoch_transport_flag", False))
    return ip, rp


class RefRun:
    """One reference run: setup, step, read back."""

    _active = None

    def __init__(self, cfg, B, flavour: str = "ieee"):
        self.lib = _lib(flavour)
        if RefRun._active is not None and RefRun._active.get(flavour) is not None:
            RefRun._active[flavour].close()
        RefRun._active = RefRun._active or {}
        self.cfg = cfg
        self.flavour = flavour
        ip, rp = pack_params(cfg)
        bdem = np.ascontiguousarray(B, dtype=np.float64).reshape(-1)
        rc = self.lib.imexref_setup(_ip(ip), _dp(rp), _dp(bdem))
        if rc != 0:
            raise RuntimeError(f"imexref_setup failed: {rc}")
        self.nv = self.lib.imexref_n_vars()
        self.open = True
        RefRun._active[flavour] = self

    def close(self):
        if self.open:
            self.lib.imexref_teardown()
            self.open = False
            if RefRun._active.get(self.flavour) is self:
                RefRun._active[self.flavour] = None

    def run(self, nsteps: int):
        dts = np.zeros(max(nsteps, 1), np.float64)
        nc = np.zeros(max(nsteps, 1), np.int32)
        rc = self.lib.imexref_run(ctypes.c_int(nsteps), _dp(dts), _ip(nc))
        if rc != 0:
            raise RuntimeError(f"imexref_run failed: {rc}")
        return dts[:nsteps], nc[:nsteps]

    def set_time(self, t, dt, dt_old, dt_old_old):
        self.lib.imexref_set_time(ctypes.c_double(t), ctypes.c_double(dt),
                                  ctypes.c_double(dt_old), ctypes.c_double(dt_old_old))

    def q(self):
        """Conservative state as [k, j, var] (C order of the Fortran q(var,j,k))."""
        out = np.zeros((self.cfg.ny, self.cfg.nx, self.nv), np.float64)
        self.lib.imexref_get_q(_dp(out))
        return out

    def set_q(self, q):
        a = np.ascontiguousarray(q, dtype=np.float64)
        assert a.shape == (self.cfg.ny, self.cfg.nx, self.nv)
        self.lib.imexref_set_q(_dp(a))

    def set_erodible(self, e):
        """erodible as [class, k, j]."""
        a = np.ascontiguousarray(np.transpose(np.asarray(e, dtype=np.float64), (1, 2, 0)))
        assert a.shape == (self.cfg.ny, self.cfg.nx, self.cfg.n_solid)
        self.lib.imexref_set_erodible(_dp(a))

    def set_noise(self, Z):
        """solver_2d's Z as [k, j]."""
        a = np.ascontiguousarray(Z, dtype=np.float64)
        assert a.shape == (self.cfg.ny, self.cfg.nx)
        self.lib.imexref_set_noise(_dp(a))

    def topography(self):
        """B_cent as [k, j]."""
        b = np.zeros((self.cfg.ny, self.cfg.nx))
        self.lib.imexref_get_bcent(_dp(b))
        return b

    def erosion(self):
        """(erosion [class, k, j], erodible [class, k, j])."""
        c = self.cfg
        ers = np.zeros((c.n_solid, c.ny, c.nx))
        erd = np.zeros((c.ny, c.nx, c.n_solid))
        self.lib.imexref_get_erosion(_dp(ers), _dp(erd))
        return ers, np.transpose(erd, (2, 0, 1)).copy()

    def qp(self):
        """The driver's module array qp as [k, j, var] (n_vars + 2 variables)."""
        out = np.zeros((self.cfg.ny, self.cfg.nx, self.nv + 2), np.float64)
        self.lib.imexref_get_qp(_dp(out))
        return out

    def interp(self, xy, f):
        """interp_2d_scalarB of the field f ([k, j]) at the points xy ([n, 2])."""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        f = np.ascontiguousarray(f, dtype=np.float64)
        assert f.shape == (self.cfg.ny, self.cfg.nx)
        out = np.zeros(len(xy), np.float64)
        self.lib.imexref_interp(ctypes.c_int(len(xy)), _dp(xy), _dp(f), _dp(out))
        return out

    def fields(self):
        c = self.cfg
        hm = np.zeros((c.ny, c.nx))
        pd = np.zeros((c.ny, c.nx))
        mv = np.zeros((c.ny, c.nx))
        dep = np.zeros((c.n_solid, c.ny, c.nx))
        t = ctypes.c_double()
        dt = ctypes.c_double()
        self.lib.imexref_get_fields(_dp(hm), _dp(pd), _dp(mv), _dp(dep),
                                    ctypes.byref(t), ctypes.byref(dt))
        return dict(hmax=hm, pdynmax=pd, mod_vel_max=mv, deposit=dep,
                    t=t.value, dt=dt.value)

    def vuln(self):
        c = self.cfg
        nl = max(1, len(c.thickness_levels) * len(c.dyn_pres_levels))
        v = np.zeros((c.ny, c.nx, nl), np.int32)
        self.lib.imexref_get_vuln(_ip(v))
        return v

    def mask(self):
        c = self.cfg
        m = np.zeros((c.ny, c.nx), np.int32)
        self.lib.imexref_get_mask(_ip(m))
        return m

    def static(self):
        c = self.cfg
        bx = np.zeros((c.ny, c.nx))
        by = np.zeros((c.ny, c.nx))
        bc = np.zeros((c.ny, c.nx))
        sc = np.zeros((c.ny, c.nx), np.int32)
        sew = np.zeros((c.ny, c.nx), np.int32)
        ap = np.zeros((c.ny, c.nx))
        ax = np.zeros((c.ny, c.nx))
        ay = np.zeros((c.ny, c.nx))
        hs = ctypes.c_double()
        self.lib.imexref_get_static(_dp(bx), _dp(by), _dp(bc), _ip(sc), _ip(sew),
                                    _dp(ap), _dp(ax), _dp(ay), ctypes.byref(hs))
        return dict(B_prime_x=bx, B_prime_y=by, B_cent=bc, source_cell=sc,
                    source_sides=sew, arc_perim=ap, arc_n_x=ax, arc_n_y=ay,
                    h_source=hs.value)

    # ---- per-function entry points (state arrays are [n, nvars]) ----------
    def qc_to_qp(self, qc):
        qc = np.ascontiguousarray(qc, np.float64)
        n = qc.shape[0]
        qp = np.zeros((n, self.nv + 2))
        pd = np.zeros(n)
        self.lib.imexref_qc_to_qp(ctypes.c_int(n), _dp(qc), _dp(qp), _dp(pd))
        return qp, pd

    def qp_to_qc(self, qp):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        qc = np.zeros((n, self.nv))
        self.lib.imexref_qp_to_qc(ctypes.c_int(n), _dp(qp), _dp(qc))
        return qc

    def mixt_var(self, qp):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, 6))
        self.lib.imexref_mixt_var(ctypes.c_int(n), _dp(qp), _dp(out))
        return out

    def fluxes(self, qc, qp, direction):
        qc = np.ascontiguousarray(qc, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.imexref_fluxes(ctypes.c_int(n), ctypes.c_int(direction), _dp(qc),
                                _dp(qp), _dp(out))
        return out

    def speeds(self, qp, direction):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, 2))
        self.lib.imexref_speeds(ctypes.c_int(n), ctypes.c_int(direction), _dp(qp),
                                _dp(out))
        return out

    def implicit(self, qc):
        qc = np.ascontiguousarray(qc, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.imexref_implicit(ctypes.c_int(n), _dp(qc), _dp(out))
        return out

    def implicit_cplx(self, qre, qim):
        qre = np.ascontiguousarray(qre, np.float64)
        qim = np.ascontiguousarray(qim, np.float64)
        n = qre.shape[0]
        ore = np.zeros((n, self.nv))
        oim = np.zeros((n, self.nv))
        self.lib.imexref_implicit_cplx(ctypes.c_int(n), _dp(qre), _dp(qim),
                                       _dp(ore), _dp(oim))
        return ore, oim

    def semi_impl(self, qc, qp):
        qc = np.ascontiguousarray(qc, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qc.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.imexref_semi_impl(ctypes.c_int(n), _dp(qc), _dp(qp), _dp(out))
        return out

    def expl(self, bx, by, t, qp):
        bx = np.ascontiguousarray(bx, np.float64)
        by = np.ascontiguousarray(by, np.float64)
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        out = np.zeros((n, self.nv))
        self.lib.imexref_expl(ctypes.c_int(n), _dp(bx), _dp(by), ctypes.c_double(t),
                              _dp(qp), _dp(out))
        return out

    def settling(self, diam, rhos, rhoc, ikv):
        a = [np.ascontiguousarray(x, np.float64) for x in (diam, rhos, rhoc, ikv)]
        n = a[0].shape[0]
        out = np.zeros(n)
        self.lib.imexref_settling(ctypes.c_int(n), *[_dp(x) for x in a], _dp(out))
        return out

    def newton(self, dt, q_si, q0, R):
        """solve_rk_step at the stage-2 coefficients, per row of q_si."""
        a = [np.ascontiguousarray(np.atleast_2d(x), np.float64) for x in (q_si, q0, R)]
        out = np.zeros_like(a[0])
        self.lib.imexref_newton(ctypes.c_int(len(a[0])), ctypes.c_double(dt), _dp(a[0]),
                                _dp(a[1]), _dp(a[2]), _dp(out))
        return out

    def mass_exchange(self, qp, dt):
        qp = np.ascontiguousarray(qp, np.float64)
        n = qp.shape[0]
        dep = np.zeros((n, self.cfg.n_solid))
        eq = np.zeros((n, self.nv))
        self.lib.imexref_mass_exchange(ctypes.c_int(n), _dp(qp), ctypes.c_double(dt),
                                       _dp(dep), _dp(eq))
        return dep, eq


