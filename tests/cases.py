"""Small parity cases shared by the tests and by tests/golden/make_golden.py.

Each returns (cfg, B) of a CPU-sized run with the physics of one of the
synthetic workloads.
"""
from __future__ import annotations

import numpy as np

from imex_sflow2d_dev_amd import workloads


def c1_48():
    """Collapse on the slope-with-bumps bed, 48 x 48 cells of 10 m."""
    cfg, B = workloads.config1(nx=48, ny=48, dx=10.0)
    cfg.x_collapse, cfg.y_collapse, cfg.r_collapse = 150.0, 240.0, 90.0
    return cfg, B


def c1_64():
    """Collapse on the slope-with-bumps bed, 64 x 64 cells of 10 m."""
    cfg, B = workloads.config1(nx=64, ny=64, dx=10.0)
    cfg.x_collapse, cfg.y_collapse, cfg.r_collapse = 200.0, 320.0, 100.0
    return cfg, B


def c2_64():
    """Radial source on the fractal cone, two classes, settling and
    entrainment, 64 x 64 cells of 20 m."""
    return workloads.small_config2(nx=64, ny=64, dx=20.0)


def c1d_96():
    """A collapse of two solid classes (1 mm at 2500 kg/m3, 0.1 mm at
    2000 kg/m3, fractions 0.03 / 0.02) with settling, in the middle of the
    slope-with-bumps bed on 96 x 96 cells of 10 m: nothing reaches the
    boundary, so each class's mass stays in the flow or in the deposit
    until buoyant cells lift off."""
    cfg, B = workloads.config1(nx=96, ny=96, dx=10.0)
    cfg.x_collapse, cfg.y_collapse, cfg.r_collapse = 480.0, 480.0, 100.0
    cfg.n_solid = 2
    cfg.rho_s, cfg.diam_s, cfg.sp_heat_s = [2500.0, 2000.0], [1e-3, 1e-4], [1200.0, 1100.0]
    cfg.alphas_collapse, cfg.alphas_source = [0.03, 0.02], [0.0, 0.0]
    cfg.settling_flag = True
    cfg.validate()
    return cfg, B


def c3_64():
    """The three-class form of the collapse (classes of 1e-2, 1e-3 and
    1e-4 m at fractions 0.02/0.02/0.01) with settling, 64 x 64 cells of
    10 m: every kernel instantiated for three classes, and a deposit."""
    cfg, B = workloads.config3(nx=64, ny=64, dx=10.0)
    cfg.x_collapse, cfg.y_collapse, cfg.r_collapse = 200.0, 320.0, 100.0
    cfg.settling_flag = True
    cfg.validate()
    return cfg, B


def c2_96():
    """As c2_64 on 96 x 96 cells."""
    return workloads.small_config2(nx=96, ny=96, dx=20.0)


def _erosion(cfg, fract):
    cfg.erosion_coeff = 2.0e-4
    cfg.erodible_porosity = 0.3
    cfg.erodible_fract = list(fract)
    cfg.T_erodible = 350.0
    cfg.erodible_thickness = 0.4
    return cfg


def c1e_48():
    """c1_48 over an erodible layer of 0.4 m (porosity 0.3, 350 K)."""
    cfg, B = c1_48()
    return _erosion(cfg, [1.0]), B


def c2e_64():
    """c2_64 over an erodible layer of 0.4 m, fractions 0.7 / 0.3."""
    cfg, B = c2_64()
    return _erosion(cfg, [0.7, 0.3]), B


def c1m_64():
    """c1_64 with the mu(I) rheology (model 11) and 5 cm particles, so the
    inertial number spans the regularised range."""
    cfg, B = c1_64()
    cfg.rheology_model = 11
    cfg.mu_s, cfg.mu_2, cfg.muI_inf, cfg.I_0 = 0.12, 0.45, 0.05, 0.279
    cfg.diam_s = [0.05]
    return cfg, B


def c2m_64():
    """c2_64 with mu(I) plus the Voellmy turbulent term (model 12); two
    classes, so the Sauter diameter and the mean particle density enter."""
    cfg, B = c2_64()
    cfg.rheology_model = 12
    cfg.mu_s, cfg.mu_2, cfg.muI_inf, cfg.I_0 = 0.08, 0.3, 0.05, 0.279
    return cfg, B


def stochastic_noise(cfg):
    """A sample of the stochastic values for a case: the host burn-in with
    std_max 600 around xi = 1000 (so a few per cent of the cells take the
    lower limit xi_temp = 1), plus cells planted beyond both limits."""
    from imex_sflow2d_dev_amd import stochastic
    cfg.std_max, cfg.tau_stochastic = 600.0, 1.0
    Z = stochastic.steady_state_noise(cfg, seed=4)
    Z[::7, ::5] = 2.5e5
    Z[3::11, 2::9] = -4000.0
    return Z


def c1s_64():
    """c1_64 with the stochastic friction (Voellmy xi shifted per cell)."""
    cfg, B = c1_64()
    cfg.stochastic_flag = True
    return cfg, B


def c2s_64():
    """c2m_64 (model 12) with the stochastic friction."""
    cfg, B = c2m_64()
    cfg.stochastic_flag = True
    return cfg, B


def _pore(cfg, B, mode, dyn, p_frac):
    """cfg with the pore-pressure physics of the TEST_DAMBREAK decks
    (TESTS/TEST_DAMBREAK/TEST_*_FINH/IMEX_SfloW2D.inp: mu(I) with mu_s 0.55,
    mu_2 0.73, muI_inf 0.05, I_0 0.279; 80 um particles of 2500 kg/m3;
    hydraulic permeability 1.1e-11 m2, gas loss, alpha_trans 0.4, maximum
    packing 0.64), and the initial state with an excess pore pressure of
    p_frac times the weight of the column (the decks start close to it:
    their restart state has p_exc = 0.9995 q1 g)."""
    from imex_sflow2d_dev_amd import init
    cfg.pore_pressure_flag = True
    cfg.gas_loss_flag = True
    cfg.f_inhibit_mode = mode
    cfg.dynamic_permeability_flag = dyn
    cfg.hydraulic_permeability = 1.1e-11
    cfg.alpha_trans = 0.4
    cfg.maximum_solid_packing = 0.64
    cfg.validate()
    st, q0 = init.setup(cfg, B)
    ns = cfg.n_solid
    q0[..., 4 + ns] = np.where(q0[..., 0] > 0.0, p_frac * cfg.grav * q0[..., 0] * q0[..., 0], 0.0)
    return cfg, B, st, q0


def _dambreak_bed(nx, ny, dx):
    """A cylinder of dense granular flow (solid fraction 0.5, 0.36 m thick,
    300 K, at rest) on a flat bed of 1 cm cells, the decks' scale."""
    cfg, B = workloads.config1(nx=nx, ny=ny, dx=dx)
    B = np.zeros_like(B)
    cfg.x_collapse, cfg.y_collapse = 0.5 * nx * dx, 0.5 * ny * dx
    cfg.r_collapse, cfg.h_collapse, cfg.T_collapse = 0.15, 0.36, 300.0
    cfg.rheology_model = 11
    cfg.mu_s, cfg.mu_2, cfg.muI_inf, cfg.I_0 = 0.55, 0.73, 0.05, 0.279
    cfg.rho_s, cfg.diam_s, cfg.sp_heat_s = [2500.0], [80e-6], [1617.0]
    cfg.alphas_collapse = [0.5]
    cfg.sp_heat_a, cfg.sp_gas_const_a, cfg.kin_visc_a = 998.0, 287.051, 1.48e-5
    cfg.pres, cfg.T_ambient = 101300.0, 300.0
    cfg.dt0, cfg.max_dt = 1e-5, 1e-2
    return cfg, B


def cp_64():
    """TEST_STATIC_FINH's physics at 64 x 64 cells of 1 cm: STATIC
    f_inhibit, dynamic (Carman-Kozeny) permeability, gas loss, mu(I);
    p_exc starts at 0.9 of the column weight.  Returns (cfg, B, st, q0)."""
    cfg, B = _dambreak_bed(64, 64, 0.01)
    return _pore(cfg, B, "STATIC", True, 0.9)


def cpd_64():
    """TEST_DYNAMIC_FINH's physics (f_inhibit's transition fraction
    from the thickness) at 64 x 64 cells of 1 cm."""
    cfg, B = _dambreak_bed(64, 64, 0.01)
    return _pore(cfg, B, "DYNAMIC", True, 0.9)


def cpo_64():
    """TEST_OFF_FINH's f_inhibit with a constant permeability and the
    Voellmy friction (model 1, whose Coulomb part takes the pore pressure
    too), mu 0.3, on 64 x 64 cells of 1 cm."""
    cfg, B = _dambreak_bed(64, 64, 0.01)
    cfg.rheology_model = 1
    cfg.mu, cfg.xi = 0.3, 1000.0
    return _pore(cfg, B, "OFF", False, 0.6)


def cps_64():
    """TEST_STEP_FINH's f_inhibit (a step at the maximum packing) with
    two classes (the Sauter diameter and the surface-weighted sphericity of
    the permeability) at 64 x 64 cells of 1 cm."""
    cfg, B = _dambreak_bed(64, 64, 0.01)
    cfg.n_solid = 2
    cfg.rho_s, cfg.diam_s, cfg.sp_heat_s = [2500.0, 2000.0], [80e-6, 300e-6], [1617.0, 1400.0]
    cfg.alphas_collapse, cfg.alphas_source = [0.35, 0.2], [0.0, 0.0]
    cfg.sphericity_s = [0.8, 0.6, 1.0]
    return _pore(cfg, B, "STEP", True, 0.9)


PORE_CASES = {"cp_64": cp_64, "cpd_64": cpd_64, "cpo_64": cpo_64, "cps_64": cps_64}


def with_transported_noise(cfg, q0, Z):
    """cfg and q0 with the noise transported by the flow
    (stoch_transport_flag): the conserved variable q1 * Z inserted after the
    solid classes, as getSteadyStateZ sets it
    (src/stochastic_module.f90:110-126)."""
    from imex_sflow2d_dev_amd import stochastic
    cfg.stochastic_flag = True
    cfg.stoch_transport_flag = True
    cfg.validate()
    q = np.insert(q0, 4 + cfg.n_solid, 0.0, axis=-1)
    return cfg, stochastic.transport_noise(cfg, q, Z)


def c1t_64():
    """c1s_64 (collapse, Voellmy friction with the stochastic xi) with the
    noise transported: (cfg, B, static, q0, Z)."""
    from imex_sflow2d_dev_amd import init
    cfg, B = c1_64()
    cfg.stochastic_flag = True
    Z = stochastic_noise(cfg)
    st, q0 = init.setup(cfg, B)
    cfg, q0 = with_transported_noise(cfg, q0, Z)
    return cfg, B, st, q0, Z


def cpot_64():
    """cpo_64 (pore pressure, f_inhibit OFF, Voellmy friction, gas loss)
    with the stochastic xi and the noise transported: the gas loss makes the
    noise's mass-exchange source eqns_term(1) * Z non-zero."""
    cfg, B, st, q0 = cpo_64()
    Z = stochastic_noise(cfg)
    cfg, q0 = with_transported_noise(cfg, q0, Z)
    return cfg, B, st, q0, Z


def c2t_48():
    """Two classes, the transported noise and no pore pressure (the general
    path's 2 x 2 Newton system) on 48 x 48 cells of 10 m."""
    from imex_sflow2d_dev_amd import init
    cfg, B = workloads.config1(nx=48, ny=48, dx=10.0)
    cfg.x_collapse, cfg.y_collapse, cfg.r_collapse = 150.0, 240.0, 90.0
    cfg.n_solid = 2
    cfg.rho_s, cfg.diam_s, cfg.sp_heat_s = [2500.0, 1500.0], [1e-3, 1e-4], [1200.0, 1200.0]
    cfg.alphas_collapse, cfg.alphas_source = [0.06, 0.04], [0.0, 0.0]
    cfg.stochastic_flag = True
    Z = stochastic_noise(cfg)
    st, q0 = init.setup(cfg, B)
    cfg, q0 = with_transported_noise(cfg, q0, Z)
    return cfg, B, st, q0, Z


TRANSPORT_CASES = {"c1t_64": c1t_64, "cpot_64": cpot_64, "c2t_48": c2t_48}


def dambreak_deck(mode="STATIC"):
    """The reference's TEST_DAMBREAK decks (TESTS/TEST_DAMBREAK/TEST_<mode>_FINH/
    IMEX_SfloW2D.inp, f_inhibit_mode <mode>): 300 x 15 cells of 1 cm, flat
    bed, mu(I) friction, one 80 um class, pore pressure with gas loss and
    dynamic permeability, per-variable boundaries (h and h alpha_s one-sided
    slope, flag 2; the others zero slope, flag 1), the deck's restart state
    example_2D_0000.q_2d (committed as tests/golden/dambreak_q0.npz), to
    T_END = 0.3 s.  Returns (cfg, B, st, q0)."""
    from conftest import golden_path
    from imex_sflow2d_dev_amd import init
    from imex_sflow2d_dev_amd.config import RunConfig
    cfg = RunConfig(nx=300, ny=15, dx=0.01, x0=0.0, y0=0.0)
    cfg.t_start, cfg.t_end = 0.0, 0.3
    cfg.dt0, cfg.max_dt, cfg.cfl = 1.0e-5, 1.0e-2, 0.23999999463558197
    cfg.n_solid, cfg.energy_flag = 1, True
    cfg.grav = 9.8100000000000005
    cfg.rheology_model = 11
    cfg.mu_s, cfg.mu_2, cfg.muI_inf, cfg.I_0 = 0.55, 0.73, 0.05, 0.279
    cfg.rho_s, cfg.diam_s, cfg.sp_heat_s = [2500.0], [80.0e-6], [1617.0]
    cfg.alphas_collapse, cfg.alphas_source = [0.0], [0.0]
    cfg.collapsing_volume_flag = False
    cfg.sp_heat_a, cfg.sp_gas_const_a, cfg.kin_visc_a = 998.0, 287.051, 1.48e-5
    cfg.pres, cfg.T_ambient = 101300.0, 300.0
    cfg.pore_pressure_flag, cfg.gas_loss_flag = True, True
    cfg.hydraulic_permeability, cfg.alpha_trans = 1.10e-11, 0.4
    cfg.f_inhibit_mode, cfg.dynamic_permeability_flag = mode, True
    cfg.maximum_solid_packing = 0.640
    side = [2, 1, 1, 1, 2, 1]
    cfg.bc_flags = [list(side) for _ in range(4)]
    cfg.bc_values = [[0.0] * 6 for _ in range(4)]
    cfg.validate()
    B = np.zeros((cfg.ny, cfg.nx))
    st, _ = init.setup(cfg, B)
    q0 = np.load(golden_path("dambreak_q0.npz"), allow_pickle=False)["q0"]
    return cfg, B, st, q0


CASES = {"c1_48": c1_48, "c1_64": c1_64, "c1d_96": c1d_96, "c3_64": c3_64, "c2_64": c2_64, "c2_96": c2_96,
         "c1e_48": c1e_48, "c2e_64": c2e_64, "c1m_64": c1m_64, "c2m_64": c2m_64,
         "c1s_64": c1s_64, "c2s_64": c2s_64}
