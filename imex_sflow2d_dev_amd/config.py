"""Run configuration for the IMEX-RK stage of the gas-particle shallow-water
system (gas carrier, Voellmy or mu(I) friction, KT fluxes, minmod MUSCL,
IMEX RK2).

The fields are the namelist values the reference reads for this path
(src/inpout_2d.f90:326-391 namelists; defaults at :408-970) plus the derived
constants read_param computes (src/inpout_2d.f90:1356 inv_pres, :1379
rho_a_amb, :2045-2049 n_vars/idx_u/idx_v, :2070-2080 inv_rho_s,
:2271-2272 limiter of u and v).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np


@dataclasses.dataclass
class RunConfig:
    # grid (RUN_PARAMETERS / DEM header)
    nx: int = 256
    ny: int = 256
    dx: float = 10.0                  # cell_size; dx = dy (src/geometry_2d.f90:272-284)
    x0: float = 0.0
    y0: float = 0.0
    # numerics (NUMERIC_PARAMETERS)
    cfl: float = 0.24
    dt0: float = 1.0e-4
    max_dt: float = 1.0
    limiter: int = 1                  # 1 = minmod (src/geometry_2d.f90:1803-1806)
    bc_flag: int = 1                  # 1 = Neumann zero slope on all sides
    t_start: float = 0.0
    t_end: float = 1.0e6              # no clipping of dt by the end time
    # physics (EXPL_TERMS / GAS / SOLID / RHEOLOGY parameters)
    n_solid: int = 1
    energy_flag: bool = True
    settling_flag: bool = False
    entrainment_flag: bool = False
    grav: float = 9.81
    mu: float = 0.1
    xi: float = 1000.0
    pres: float = 101325.0
    T_ambient: float = 300.0
    sp_heat_a: float = 1004.0
    sp_gas_const_a: float = 287.0
    kin_visc_a: float = 1.48e-5
    rho_s: List[float] = dataclasses.field(default_factory=lambda: [2500.0])
    diam_s: List[float] = dataclasses.field(default_factory=lambda: [1.0e-4])
    sp_heat_s: List[float] = dataclasses.field(default_factory=lambda: [1200.0])
    # initial collapsing volume (COLLAPSING_VOLUME_PARAMETERS)
    collapsing_volume_flag: bool = True
    x_collapse: float = 640.0
    y_collapse: float = 1280.0
    r_collapse: float = 300.0
    h_collapse: float = 20.0
    T_collapse: float = 600.0
    alphas_collapse: List[float] = dataclasses.field(default_factory=lambda: [0.05])
    # radial lateral source (RADIAL_SOURCE_PARAMETERS)
    radial_source_flag: bool = False
    x_source: float = 0.0
    y_source: float = 0.0
    r_source: float = 0.0
    vel_source: float = 0.0
    T_source: float = 0.0
    mfr_source: float = -1.0
    h_source: float = 0.0             # derived from mfr_source when > 0
    time_param: List[float] = dataclasses.field(default_factory=lambda: [0.0, 0.0, 0.0, 0.0])
    alphas_source: List[float] = dataclasses.field(default_factory=lambda: [0.0])
    # erosion of an erodible layer (SOLID_TRANSPORT_PARAMETERS,
    # src/inpout_2d.f90:1813-1900); the layer's thickness per class is
    # erodible_fract * (1 - erodible_porosity) * erodible_thickness
    # (src/inpout_2d.f90:1982-1986), uniform here
    erosion_coeff: float = 0.0
    erodible_porosity: float = 0.0
    erodible_fract: List[float] = dataclasses.field(default_factory=list)
    T_erodible: float = 300.0
    erodible_deposit_flag: bool = False
    erodible_thickness: float = 0.0
    # B_cent follows erosion and deposition (topo_change_flag), with the bed
    # slopes recomputed every step (src/IMEX_SfloW2D.f90:445)
    topo_change_flag: bool = False
    # friction (RHEOLOGY_PARAMETERS; mu(I) defaults -1 at src/inpout_2d.f90:511-514,
    # checks at :3592-3630):
    # 1 = Voellmy-Salm (mu, xi); 11 = mu(I) for dense granular flows
    # (mu_s, mu_2, muI_inf, I_0); 12 = mu(I) plus the Voellmy turbulent term
    rheology_model: int = 1
    mu_s: float = -1.0
    mu_2: float = -1.0
    muI_inf: float = -1.0
    I_0: float = -1.0
    # stochastic friction (STOCHASTIC_PARAMETERS, defaults src/inpout_2d.f90:
    # 526-533): a per-cell value Z shifts xi (models 1 and 12); the noise is
    # not transported.  Z is 0 unless given (the reference's value outside
    # an ASCII restart); stochastic.steady_state_noise gives the burn-in of
    # getSteadyStateZ from std_max, tau_stochastic, sym_noise, noise_pow_val
    # and length_spatial_corr
    stochastic_flag: bool = False
    std_max: float = -1.0
    tau_stochastic: float = -1.0
    sym_noise: float = 0.0
    noise_pow_val: float = -1.0
    length_spatial_corr: float = -1.0
    # the transported noise (stoch_transport_flag, src/inpout_2d.f90:1466-1475):
    # q1 * Z as a conserved variable after the solid classes, advected with
    # the flow; its initial value is q1 times the noise (getSteadyStateZ,
    # src/stochastic_module.f90:110-126), the friction keeps reading the
    # fixed noise
    stoch_transport_flag: bool = False
    # pore pressure (PORE_PRESSURE_PARAMETERS, src/inpout_2d.f90:389-391,
    # defaults :536-542; maximum_solid_packing and sphericity_s from
    # SOLID_TRANSPORT_PARAMETERS, defaults :819, :854): the conserved
    # variable q1 * p_exc after the solid classes, its implicit decay with
    # f_inhibit (OFF, STEP, STATIC, DYNAMIC) and the gas loss
    pore_pressure_flag: bool = False
    gas_loss_flag: bool = False
    f_inhibit_mode: str = "OFF"
    N_inh: int = 2
    alpha_trans: float = 0.0
    hydraulic_permeability: float = 0.0
    dynamic_permeability_flag: bool = False
    maximum_solid_packing: float = 0.64
    sphericity_s: List[float] = dataclasses.field(default_factory=lambda: [1.0, 1.0, 1.0])
    # per-variable boundary conditions (WEST/EAST/SOUTH/NORTH_BOUNDARY_
    # CONDITIONS, src/solver_2d.f90:3190-3484): bc_flags[side][var] with
    # sides W, E, S, N and the conservative variables' order; flag 0
    # Dirichlet value, 1 slope value, 2 one-sided slope.  None: flag 1,
    # value 0 everywhere (the path's default; other boundaries are on the
    # pore-pressure path only)
    bc_flags: Optional[List[List[int]]] = None
    bc_values: Optional[List[List[float]]] = None
    # vulnerability table (VULNERABILITY_TABLE_PARAMETERS)
    thickness_levels: List[float] = dataclasses.field(default_factory=list)
    dyn_pres_levels: List[float] = dataclasses.field(default_factory=list)
    name: str = "run"
    # 0: FP64 drop-in; 1: FP32 form (state, reconstruction and fluxes in FP32)
    precision: int = 0

    # ---- derived quantities ------------------------------------------------
    @property
    def n_vars(self) -> int:
        return (4 + self.n_solid + (1 if self.stoch_transport_flag else 0)
                + (1 if self.pore_pressure_flag else 0))

    @property
    def f_inhibit_code(self) -> int:
        return ("OFF", "STEP", "STATIC", "DYNAMIC").index(self.f_inhibit_mode)

    @property
    def n_eqns(self) -> int:
        return self.n_vars

    @property
    def inv_pres(self) -> float:
        return 1.0 / self.pres

    @property
    def rho_a_amb(self) -> float:
        return self.pres / (self.sp_gas_const_a * self.T_ambient)

    @property
    def eps_sing(self) -> float:
        # src/geometry_2d.f90:300-310: MIN(MIN(dx**4, dy**4), 1e-10)
        d4 = math.pow(self.dx, 4.0)
        return min(min(d4, d4), 1.0e-10)

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny

    def erodible_fractions(self) -> List[float]:
        """erodible_fract as read_param leaves it (src/inpout_2d.f90:1827-1866):
        1/n_solid without erosion, 1 for one class, else the given fractions
        divided by their sum when it is not one."""
        ns = self.n_solid
        if self.erosion_coeff == 0.0:
            return [1.0 / ns] * ns
        if ns == 1:
            return [1.0]
        f = [float(x) for x in self.erodible_fract[:ns]]
        s = sum(f)
        return [x / s for x in f] if s != 1.0 else f

    def effective_T_erodible(self) -> float:
        """T_erodible: 300 K without erosion (src/inpout_2d.f90:1831)."""
        return 300.0 if self.erosion_coeff == 0.0 else self.T_erodible

    def effective_erodible_porosity(self) -> float:
        return 0.0 if self.erosion_coeff == 0.0 else self.erodible_porosity

    def validate(self) -> None:
        if self.nx < 5 or self.ny < 5:
            raise ValueError("the 5-point bed-slope stencil needs nx, ny >= 5")
        if not (1 <= self.n_solid <= 3):
            raise ValueError("n_solid must be in 1..3")
        for name in ("rho_s", "diam_s", "sp_heat_s", "alphas_collapse", "alphas_source"):
            if len(getattr(self, name)) < self.n_solid:
                raise ValueError(f"{name} needs n_solid entries")
        if self.cfl > 0.25:
            # src/inpout_2d.f90:2261-2267 pauses for cfl > 0.25 in 2-D
            raise ValueError("cfl must be <= 0.25 for a 2-D run")
        if len(self.thickness_levels) > 8 or len(self.dyn_pres_levels) > 8:
            raise ValueError("at most 8 thickness and 8 dynamic-pressure levels")
        if self.limiter != 1:
            raise ValueError("only the minmod limiter (1) is on this path")
        if self.bc_flag != 1:
            raise ValueError("only Neumann (flag 1) boundaries are on this path")
        if self.bc_flags is not None:
            if not (self.pore_pressure_flag or self.stoch_transport_flag):
                raise ValueError("per-variable boundary conditions are on the pore-pressure and "
                                 "transported-noise path only")
            if (len(self.bc_flags) != 4 or len(self.bc_values) != 4
                    or any(len(f) < self.n_vars for f in self.bc_flags)
                    or any(len(v) < self.n_vars for v in self.bc_values)
                    or any(x not in (0, 1, 2) for f in self.bc_flags for x in f[:self.n_vars])):
                raise ValueError("bc_flags / bc_values: 4 sides x n_vars, flags 0, 1 or 2")
        if self.rheology_model not in (1, 11, 12):
            raise ValueError("rheology_model must be 1 (Voellmy-Salm), 11 or 12 (mu(I))")
        if self.rheology_model in (11, 12) and min(self.mu_s, self.mu_2, self.I_0,
                                                    self.muI_inf) < 0.0:
            # src/inpout_2d.f90:3600-3611, :3622-3629
            raise ValueError("mu(I) rheology needs mu_s, mu_2, I_0 and muI_inf >= 0")
        if self.erosion_coeff < 0.0:
            raise ValueError("erosion_coeff must be >= 0")
        if self.f_inhibit_mode not in ("OFF", "STEP", "STATIC", "DYNAMIC"):
            # src/inpout_2d.f90:681-687
            raise ValueError("f_inhibit_mode must be OFF, STEP, STATIC or DYNAMIC")
        if not (0.0 < self.maximum_solid_packing < 1.0):
            raise ValueError("maximum_solid_packing must be in (0, 1)")  # :1989-1995
        if any(not (0.0 < x <= 1.0) for x in self.sphericity_s[:self.n_solid]):
            raise ValueError("sphericity_s must be in (0, 1]")  # :1797-1803
        if self.stoch_transport_flag:
            if not self.stochastic_flag:
                raise ValueError("stoch_transport_flag needs stochastic_flag")
            if self.radial_source_flag or self.erosion_coeff > 0.0 or self.settling_flag \
                    or self.entrainment_flag:
                raise ValueError("the transported noise with the radial source, settling, "
                                 "entrainment or erosion is not on this path")
            if self.n_vars > 8:
                raise ValueError("the transported noise: at most 8 conserved variables")
        if self.pore_pressure_flag:
            if self.f_inhibit_mode in ("STATIC", "DYNAMIC"):
                if self.N_inh <= 0 or self.alpha_trans <= 0.0:
                    # src/inpout_2d.f90:698-710
                    raise ValueError("STATIC/DYNAMIC f_inhibit needs N_inh > 0 and "
                                     "alpha_trans > 0")
                if self.maximum_solid_packing < self.alpha_trans:
                    raise ValueError("maximum_solid_packing must be >= alpha_trans")
            if self.radial_source_flag:
                raise ValueError("pore pressure with the radial source is not on this path")
            if self.erosion_coeff > 0.0 or self.settling_flag or self.entrainment_flag:
                raise ValueError("pore pressure with settling, entrainment or erosion is "
                                 "not on this path")
        if self.erosion_coeff > 0.0:
            if not (0.0 <= self.erodible_porosity < 1.0):
                raise ValueError("erodible_porosity must be in [0, 1)")
            if self.n_solid > 1 and (len(self.erodible_fract) < self.n_solid
                                     or min(self.erodible_fract[:self.n_solid]) < 0.0):
                raise ValueError("erodible_fract needs n_solid non-negative entries")
            if self.T_erodible < 0.0:
                raise ValueError("T_erodible must be >= 0")


def solid_arrays(cfg: RunConfig):
    """rho_s, inv_rho_s, diam_s, sp_heat_s as float64 arrays of length n_solid."""
    ns = cfg.n_solid
    rho = np.asarray(cfg.rho_s[:ns], dtype=np.float64)
    return rho, 1.0 / rho, np.asarray(cfg.diam_s[:ns], np.float64), \
        np.asarray(cfg.sp_heat_s[:ns], np.float64)
