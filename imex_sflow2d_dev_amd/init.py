"""Initial conditions of the path (host side, numpy), restating
collapsing_volume (src/init_2d.f90:76-163) and the primitive -> conservative
conversion it uses (qp_to_qc, src/constitutive_2d.f90:1718-2149, gas carrier,
alpha_flag = F) with the reference's operation order."""
from __future__ import annotations

import numpy as np

from .config import RunConfig
from .geometry import compute_cell_fract

EPS = 2.220446049250313e-16


def qp_to_qc(cfg: RunConfig, qp: np.ndarray) -> np.ndarray:
    """Vectorised qp_to_qc: qp[..., n_vars+2] -> qc[..., n_vars]."""
    nv, ns = cfg.n_vars, cfg.n_solid
    qp = np.asarray(qp, dtype=np.float64)
    h = qp[..., 0]
    wet = h > EPS
    hs = np.where(wet, h, 1.0)
    hu, hv, T = qp[..., 1], qp[..., 2], np.where(wet, qp[..., 3], cfg.T_ambient)
    u, v = qp[..., nv], qp[..., nv + 1]
    alphas = [qp[..., 4 + i] / hs for i in range(ns)]
    at = np.zeros_like(h)
    for a in alphas:
        at = at + a
    ar = [alphas[i] * cfg.rho_s[i] for i in range(ns)]
    rho_a = cfg.pres / (cfg.sp_gas_const_a * T)
    rho_c = (((1.0 - 0.0) - at - 0.0) * rho_a + 0.0) / ((1.0 - 0.0) - at)
    alphac = 1.0 - at
    sar = np.zeros_like(h)
    for a in ar:
        sar = sar + a
    rho_m = alphac * rho_c + sar
    inv_rhom = 1.0 / rho_m
    xs = [a * inv_rhom for a in ar]
    xc = alphac * rho_c * inv_rhom
    xcs = np.where(xc > EPS, xc, 1.0)
    cp_c = np.where(xc > EPS, ((xcs - 0.0) * cfg.sp_heat_a + 0.0) / xcs, cfg.sp_heat_a)
    dot = np.zeros_like(h)
    for i in range(ns):
        dot = dot + xs[i] * cfg.sp_heat_s[i]
    cp_mix = dot + xc * cp_c
    qc = np.zeros(qp.shape[:-1] + (nv,))
    qc[..., 0] = rho_m * h
    qc[..., 1] = rho_m * hu
    qc[..., 2] = rho_m * hv
    if cfg.energy_flag:
        qc[..., 3] = np.where(h > 0.0, h * rho_m * (cp_mix * T + 0.5 * (u * u + v * v)), 0.0)
    else:
        qc[..., 3] = h * rho_m * cp_mix * T
    for i in range(ns):
        qc[..., 4 + i] = xs[i] * qc[..., 0]
    if cfg.stoch_transport_flag:
        # q1 * Z (src/constitutive_2d.f90:1979, :2141)
        qc[..., 4 + ns] = qp[..., 4 + ns] * qc[..., 0]
    if cfg.pore_pressure_flag:
        # q1 * p_exc (src/constitutive_2d.f90:1981, :2143)
        ip = 4 + ns + (1 if cfg.stoch_transport_flag else 0)
        qc[..., ip] = qp[..., ip] * qc[..., 0]
    qc[~wet] = 0.0
    return qc


def mixt_var_rho_m(cfg: RunConfig, qp: np.ndarray) -> float:
    """Mixture density of one primitive state (mixt_var, src/constitutive_2d.f90:1342-1492)."""
    ns = cfg.n_solid
    h = float(qp[0])
    if h <= EPS:
        return cfg.rho_a_amb
    T = float(qp[3])
    alphas = [float(qp[4 + i]) / h for i in range(ns)]
    at = 0.0
    for a in alphas:
        at = at + a
    rho_a = cfg.pres / (cfg.sp_gas_const_a * T)
    rc = (((1.0 - 0.0) - at - 0.0) * rho_a + 0.0) / ((1.0 - 0.0) - at)
    dot = 0.0
    for i in range(ns):
        dot = dot + alphas[i] * cfg.rho_s[i]
    return (1.0 - at) * rc + dot


def collapsing_volume(cfg: RunConfig) -> np.ndarray:
    """Initial q as [k, j, var]: a cylinder at rest, cell fractions from
    compute_cell_fract."""
    nv, ns = cfg.n_vars, cfg.n_solid
    frac = compute_cell_fract(cfg, cfg.x_collapse, cfg.y_collapse, cfg.r_collapse,
                              cfg.r_collapse, 0.0)
    qp = np.zeros((cfg.ny, cfg.nx, nv + 2))
    sel = frac > 0.0
    qp[..., 3] = cfg.T_collapse
    qp[..., 0] = np.where(sel, frac * cfg.h_collapse, 0.0)
    for i in range(ns):
        qp[..., 4 + i] = np.where(sel, frac * cfg.h_collapse * cfg.alphas_collapse[i], 0.0)
    if cfg.pore_pressure_flag:
        # the cylinder's excess pore pressure is 1 and the transported
        # noise 0 until getSteadyStateZ sets it (src/init_2d.f90:143-147)
        ip = 4 + ns + (1 if cfg.stoch_transport_flag else 0)
        qp[..., ip] = np.where(sel, 1.0, 0.0)
    return qp_to_qc(cfg, qp)


def initial_state(cfg: RunConfig) -> np.ndarray:
    if cfg.collapsing_volume_flag:
        return collapsing_volume(cfg)
    return np.zeros((cfg.ny, cfg.nx, cfg.n_vars))


def setup(cfg: RunConfig, B: np.ndarray):
    """Static fields and the initial state of a run: (static dict, q0[k, j, var])."""
    from .geometry import bed_slopes, init_source, source_thickness
    bx, by = bed_slopes(cfg, B)
    st = dict(B_prime_x=bx, B_prime_y=by, B_cent=np.asarray(B, np.float64))
    if cfg.radial_source_flag:
        sc, sd, ap, ax, ay = init_source(cfg)
        st.update(source_cell=sc, source_sides=sd, arc_perim=ap, arc_n_x=ax, arc_n_y=ay)
        if cfg.mfr_source > 0.0:
            cfg.h_source = source_thickness(cfg)
    st["h_source"] = cfg.h_source
    if cfg.erosion_coeff > 0.0:
        # erodible(i,:,:) = erodible_fract(i) * (1 - erodible_porosity) *
        # erodible_init (src/inpout_2d.f90:1982-1986), a uniform layer here
        por = cfg.effective_erodible_porosity()
        st["erodible"] = np.stack([np.full((cfg.ny, cfg.nx), f * (1.0 - por) * cfg.erodible_thickness)
                                   for f in cfg.erodible_fractions()])
    return st, initial_state(cfg)
