"""ctypes binding of libimexhip.so (the C ABI declared in include/imex_hip.h).

The library is built in-tree by imex_sflow2d_dev_amd/build.py.  There is no
fallback: if the library is missing or no GPU is present every call raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .config import RunConfig

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libimexhip.so")

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int32)

# error bits (include/imex_hip.h)
ERR_NAN = 1
ERR_NEG_H = 2
ERR_NEG_SOLID = 4
ERR_COLD = 8
ERR_NEG_DEP = 16
ERR_SOLID_GT_TOT = 32
WARN_RESCALE = 64

DL_Q, DL_STATS, DL_DEPOSIT, DL_VULN, DL_MASK = 1, 2, 4, 8, 16

EXPORTS = [
    "imex_create", "imex_set_static", "imex_upload_state", "imex_check_solve",
    "imex_timestep", "imex_rk_step", "imex_erosion_deposition", "imex_post_step",
    "imex_run_steps", "imex_run_steps_timed", "imex_calib_copy", "imex_get_time",
    "imex_get_time_history", "imex_get_counters", "imex_set_time", "imex_download", "imex_last_error", "imex_destroy",
    "imex_batch_create", "imex_batch_create_form", "imex_batch_set_blocks",
    "imex_batch_set_launch", "imex_batch_add", "imex_batch_remove",
    "imex_batch_run", "imex_batch_status", "imex_batch_last_error", "imex_batch_destroy",
    "imex_test_qc_to_qp", "imex_test_qp_to_qc", "imex_test_fluxes", "imex_test_speeds",
    "imex_test_implicit_cplx", "imex_test_implicit_real", "imex_test_semi_impl", "imex_test_expl",
    "imex_test_settling", "imex_test_mass_exchange", "imex_test_newton", "imex_test_arith",
    "imex_test_pow", "imex_host_pow",
    "imex_set_topography", "imex_runout_reduce", "imex_probes",
    "imex_snapshot_cols", "imex_snapshot_begin", "imex_snapshot_wait", "imex_host_alloc",
    "imex_host_free", "imex_write_records", "imex_write_restart", "imex_format_value",
    "imex_set_erodible", "imex_get_erosion", "imex_set_noise", "imex_test_mass_exchange_ers",
    "imex_topography_reconstruction", "imex_get_topography", "imex_get_work_counters",
    "imex_post_step_noalloc", "imex_get_footprint",
]

RUNOUT_STORE, RUNOUT_B = 1, 2
SNAP_Q, SNAP_Q2D, SNAP_DEP2D, SNAP_P2D, SNAP_FIELDS, SNAP_ERO2D, SNAP_ERD2D = 1, 2, 4, 8, 16, 32, 64
SNAP_KINDS = (SNAP_Q, SNAP_Q2D, SNAP_DEP2D, SNAP_P2D, SNAP_FIELDS, SNAP_ERO2D, SNAP_ERD2D)
FMT_E20_12, FMT_ES15_7, FMT_ES14_7 = 0, 1, 2


class ImexConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("n_solid", ctypes.c_int32),
        ("energy_flag", ctypes.c_int32), ("settling_flag", ctypes.c_int32),
        ("entrainment_flag", ctypes.c_int32), ("radial_source_flag", ctypes.c_int32),
        ("n_thickness_levels", ctypes.c_int32), ("n_dyn_pres_levels", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("dx", ctypes.c_double), ("x0", ctypes.c_double), ("y0", ctypes.c_double),
        ("cfl", ctypes.c_double), ("max_dt", ctypes.c_double),
        ("grav", ctypes.c_double), ("mu", ctypes.c_double), ("xi", ctypes.c_double),
        ("pres", ctypes.c_double), ("T_ambient", ctypes.c_double),
        ("sp_heat_a", ctypes.c_double), ("sp_gas_const_a", ctypes.c_double),
        ("kin_visc_a", ctypes.c_double),
        ("rho_s", ctypes.c_double * 3), ("diam_s", ctypes.c_double * 3),
        ("sp_heat_s", ctypes.c_double * 3),
        ("vel_source", ctypes.c_double), ("T_source", ctypes.c_double),
        ("h_source", ctypes.c_double), ("time_param", ctypes.c_double * 4),
        ("alphas_source", ctypes.c_double * 3),
        ("thickness_levels", ctypes.c_double * 8), ("dyn_pres_levels", ctypes.c_double * 8),
        ("erosion_coeff", ctypes.c_double), ("erodible_porosity", ctypes.c_double),
        ("T_erodible", ctypes.c_double), ("erodible_fract", ctypes.c_double * 3),
        ("erodible_deposit_flag", ctypes.c_int32), ("topo_change_flag", ctypes.c_int32),
        ("rheology_model", ctypes.c_int32),
        ("mu_s", ctypes.c_double), ("mu_2", ctypes.c_double), ("muI_inf", ctypes.c_double),
        ("I_0", ctypes.c_double), ("stochastic_flag", ctypes.c_int32),
        ("pore_pressure_flag", ctypes.c_int32), ("gas_loss_flag", ctypes.c_int32),
        ("f_inhibit_mode", ctypes.c_int32), ("N_inh", ctypes.c_int32),
        ("dynamic_permeability_flag", ctypes.c_int32),
        ("hydraulic_permeability", ctypes.c_double), ("alpha_trans", ctypes.c_double),
        ("maximum_solid_packing", ctypes.c_double), ("sphericity_s", ctypes.c_double * 3),
        ("bc_custom", ctypes.c_int32), ("bc_flag", ctypes.c_int32 * 32),
        ("bc_value", ctypes.c_double * 32),
        ("stoch_transport_flag", ctypes.c_int32),
    ]


class ImexStatic(ctypes.Structure):
    _fields_ = [("B_prime_x", _D), ("B_prime_y", _D), ("source_cell", _I),
                ("source_sides", _I), ("arc_perim", _D), ("arc_n_x", _D), ("arc_n_y", _D)]


class ImexStepStats(ctypes.Structure):
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("mass", ctypes.c_double),
                ("volume", ctypes.c_double), ("area", ctypes.c_double),
                ("solid_mass", ctypes.c_double), ("active_cells", ctypes.c_int64),
                ("dried_cells", ctypes.c_int64), ("err", ctypes.c_uint32)]


class ImexRunoutRed(ctypes.Structure):
    _fields_ = [("max_h", ctypes.c_double), ("sum_q1", ctypes.c_double),
                ("sum_xq1", ctypes.c_double), ("sum_yq1", ctypes.c_double),
                ("n_hpos", ctypes.c_int64), ("n_hpos_old", ctypes.c_int64),
                ("n_new", ctypes.c_int64), ("imax", ctypes.c_int64), ("imax_x", ctypes.c_int64),
                ("imax_y", ctypes.c_int64), ("dist_max", ctypes.c_double),
                ("dist_x_max", ctypes.c_double), ("dist_y_max", ctypes.c_double),
                ("dist_at_imax_x", ctypes.c_double), ("dist_at_imax_y", ctypes.c_double),
                ("ib", ctypes.c_int64), ("b_max", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


_LIB = None


class NativeError(RuntimeError):
    pass


def lib():
    """Loads libimexhip.so; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise NativeError(f"{LIB_PATH} is not built (python -m imex_sflow2d_dev_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        for name in EXPORTS:
            f = getattr(L, name)
            f.restype = None if name in ("imex_destroy", "imex_batch_destroy") else ctypes.c_int
        L.imex_create.argtypes = [ctypes.POINTER(ImexConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.imex_host_alloc.argtypes = [ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)]
        L.imex_host_free.argtypes = [ctypes.c_void_p]
        L.imex_write_records.argtypes = [ctypes.c_char_p, _D, ctypes.c_int64, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.imex_write_restart.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 4 + \
            [ctypes.c_double] * 2 + [_D] * 9
        L.imex_format_value.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_char_p]
        L.imex_get_footprint.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_size_t)]
        _LIB = L
    return _LIB


def exported_symbols():
    L = lib()
    return {n: hasattr(L, n) for n in EXPORTS}


def make_config(cfg: RunConfig) -> ImexConfig:
    c = ImexConfig()
    c.nx, c.ny, c.n_solid = cfg.nx, cfg.ny, cfg.n_solid
    c.energy_flag = int(cfg.energy_flag)
    c.settling_flag = int(cfg.settling_flag)
    c.entrainment_flag = int(cfg.entrainment_flag)
    c.radial_source_flag = int(cfg.radial_source_flag)
    c.n_thickness_levels = len(cfg.thickness_levels)
    c.n_dyn_pres_levels = len(cfg.dyn_pres_levels)
    c.precision = int(getattr(cfg, "precision", 0))
    c.dx, c.x0, c.y0 = cfg.dx, cfg.x0, cfg.y0
    c.cfl, c.max_dt = cfg.cfl, cfg.max_dt
    c.grav, c.mu, c.xi, c.pres = cfg.grav, cfg.mu, cfg.xi, cfg.pres
    c.T_ambient, c.sp_heat_a = cfg.T_ambient, cfg.sp_heat_a
    c.sp_gas_const_a, c.kin_visc_a = cfg.sp_gas_const_a, cfg.kin_visc_a
    for i in range(cfg.n_solid):
        c.rho_s[i] = cfg.rho_s[i]
        c.diam_s[i] = cfg.diam_s[i]
        c.sp_heat_s[i] = cfg.sp_heat_s[i]
        c.alphas_source[i] = cfg.alphas_source[i] if i < len(cfg.alphas_source) else 0.0
    c.vel_source, c.T_source, c.h_source = cfg.vel_source, cfg.T_source, cfg.h_source
    for i in range(4):
        c.time_param[i] = cfg.time_param[i]
    for i, v in enumerate(cfg.thickness_levels):
        c.thickness_levels[i] = v
    for i, v in enumerate(cfg.dyn_pres_levels):
        c.dyn_pres_levels[i] = v
    c.erosion_coeff = cfg.erosion_coeff
    c.erodible_porosity = cfg.effective_erodible_porosity()
    c.T_erodible = cfg.effective_T_erodible()
    for i, v in enumerate(cfg.erodible_fractions()):
        c.erodible_fract[i] = v
    c.erodible_deposit_flag = int(cfg.erodible_deposit_flag)
    c.topo_change_flag = int(cfg.topo_change_flag)
    c.rheology_model = cfg.rheology_model
    c.mu_s, c.mu_2, c.muI_inf, c.I_0 = cfg.mu_s, cfg.mu_2, cfg.muI_inf, cfg.I_0
    c.stochastic_flag = int(cfg.stochastic_flag)
    c.pore_pressure_flag = int(cfg.pore_pressure_flag)
    c.stoch_transport_flag = int(cfg.stoch_transport_flag)
    c.gas_loss_flag = int(cfg.gas_loss_flag)
    c.f_inhibit_mode = cfg.f_inhibit_code
    c.N_inh = int(cfg.N_inh)
    c.dynamic_permeability_flag = int(cfg.dynamic_permeability_flag)
    c.hydraulic_permeability = cfg.hydraulic_permeability
    c.alpha_trans = cfg.alpha_trans
    c.maximum_solid_packing = cfg.maximum_solid_packing
    for i in range(3):
        c.sphericity_s[i] = cfg.sphericity_s[i] if i < len(cfg.sphericity_s) else 1.0
    if cfg.bc_flags is not None:
        c.bc_custom = 1
        for sd in range(4):
            for i in range(cfg.n_vars):
                c.bc_flag[8 * sd + i] = int(cfg.bc_flags[sd][i])
                c.bc_value[8 * sd + i] = float(cfg.bc_values[sd][i])
    return c


def _dp(a):
    return a.ctypes.data_as(_D)


def _ip(a):
    return a.ctypes.data_as(_I)


class Handle:
    """Owns one imex_handle (device state of one run)."""

    def __init__(self, cfg: RunConfig):
        self.L = lib()
        self.cfg = cfg
        self.nv = cfg.n_vars
        self._c = make_config(cfg)
        h = ctypes.c_void_p()
        rc = self.L.imex_create(ctypes.byref(self._c), ctypes.byref(h))
        self.h = h
        if rc != 0:
            msg = self.last_error() if h.value else ""
            self.close()
            raise NativeError(f"imex_create failed ({rc}): {msg}")

    def last_error(self) -> str:
        buf = ctypes.create_string_buffer(512)
        if self.h and self.h.value:
            self.L.imex_last_error(self.h, buf, ctypes.c_size_t(512))
        return buf.value.decode(errors="replace")

    def check(self, rc: int, what: str) -> int:
        if rc < 0:
            raise NativeError(f"{what} failed ({rc}): {self.last_error()}")
        return rc

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.L.imex_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A page-locked host buffer (imex_host_alloc) viewed as a float64 array."""

    def __init__(self, n: int):
        self.L = lib()
        self.p = ctypes.c_void_p()
        if self.L.imex_host_alloc(ctypes.c_size_t(8 * max(int(n), 1)), ctypes.byref(self.p)) != 0:
            raise NativeError("imex_host_alloc failed")
        self.n = int(n)
        self.a = np.ctypeslib.as_array(ctypes.cast(self.p, _D), shape=(max(self.n, 1),))[:self.n]

    def free(self):
        if self.p is not None and self.p.value:
            self.a = None
            self.L.imex_host_free(self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def write_records(path: str, rec: np.ndarray, nx: int, fmt: int, n_threads: int = 0):
    """The reference's formatted record files (.q_2d: FMT_E20_12, .p_2d:
    FMT_ES15_7) from records [n_cells, n_col]."""
    rec = np.ascontiguousarray(rec, dtype=np.float64)
    rc = lib().imex_write_records(path.encode(), _dp(rec), ctypes.c_int64(rec.shape[0]),
                                  ctypes.c_int(rec.shape[1]), ctypes.c_int(nx), ctypes.c_int(fmt),
                                  ctypes.c_int(n_threads))
    if rc != 0:
        raise NativeError(f"imex_write_records failed ({rc}) for {path}")


def write_restart(path: str, nx, ny, n_vars, n_solid, t, dt, q_aos, B_cent, erodible, deposit,
                  erosion, hmax, pdynmax, mod_vel_max, Z=None):
    """restart.bin as write_restart_file writes it; Z (nx * ny) is the
    stochastic_flag record, omitted when None."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in
            (q_aos, B_cent, erodible, deposit, erosion, hmax, pdynmax, mod_vel_max)]
    z = None if Z is None else np.ascontiguousarray(Z, dtype=np.float64)
    rc = lib().imex_write_restart(path.encode(), nx, ny, n_vars, n_solid, t, dt,
                                  *[_dp(a) for a in arrs], None if z is None else _dp(z))
    if rc != 0:
        raise NativeError(f"imex_write_restart failed ({rc}) for {path}")


def format_value(fmt: int, x: float) -> str:
    buf = ctypes.create_string_buffer(32)
    if lib().imex_format_value(fmt, x, buf) != 0:
        raise NativeError("imex_format_value failed")
    return buf.value.decode()
