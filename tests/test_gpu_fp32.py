"""The FP32 form (state, reconstruction and fluxes in FP32; Newton, thin
layers, deposit and maxima in FP64) against the FP64 drop-in on the same
run and against the FP64 oracle: runout area, maximum thickness, deposit
and mass agree to the stated tolerance, and the FP32 run keeps the mass
budget."""
import copy

import numpy as np
import pytest

import cases
from oracle import oraclepy
from imex_sflow2d_dev_amd import init, precision, workloads
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n, dx, steps, rheology", [(96, 20.0, 400, 1), (160, 12.0, 500, 1),
                                                    (96, 20.0, 400, 11)])
def test_fp32_form_against_fp64(n, dx, steps, rheology):
    cfg, B = workloads.config3(nx=n, ny=n, dx=dx)
    if rheology == 11:                  # mu(I) friction (FP64 in both forms)
        cfg.rheology_model = 11
        cfg.mu_s, cfg.mu_2, cfg.muI_inf, cfg.I_0 = 0.12, 0.45, 0.05, 0.279
    st, q0 = init.setup(cfg, B)
    r = precision.compare(cfg, st, q0, steps)
    f64, f32 = r["fp64"], r["fp32"]
    assert f64["err_bits"] == 0 and f32["err_bits"] == 0
    assert f32["t_end_s"] == pytest.approx(f64["t_end_s"], rel=1e-3)
    rel = r["relative_difference"]
    # the stated bar of the precision form: 1e-3 relative on these quantities
    for k in ("runout_area_m2", "max_thickness_m", "deposit_volume_m3", "total_mass_kg"):
        assert rel[k] < 1e-3, (k, rel[k])


def _oracle_stats(cfg, o):
    """precision.run_stats of an oracle run (the same formulas)."""
    f = o.fields()
    q = o.q()
    area = cfg.dx * cfg.dx
    dep = f["deposit"]                       # [class, k, j]
    rho_s = np.asarray(cfg.rho_s[:cfg.n_solid])
    return {
        "runout_area_m2": float(area * np.count_nonzero(f["hmax"] > 1.0e-3)),
        "max_thickness_m": float(f["hmax"].max()),
        "deposit_volume_m3": float(area * dep.sum()),
        "total_mass_kg": float(area * q[:, :, 0].sum()
                               + float((dep.sum(axis=(1, 2)) * rho_s).sum()) * area),
    }


@pytest.mark.parametrize("steps", [150, 250])
def test_fp32_form_against_fp64_oracle_three_classes(steps):
    """The three-class collapse with settling (it deposits), 64 x 64 cells:
    the north star's bar, 1e-4 relative on runout area, maximum thickness,
    deposit and total mass, against the FP64 oracle."""
    cfg, B = cases.c3_64()
    st, q0 = init.setup(cfg, B)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    o.run(steps)
    ref = _oracle_stats(cfg, o)
    c = copy.deepcopy(cfg)
    c.precision = 1
    s = Solver(c, st, q0, raise_on_error=False)
    s.run_steps(steps)
    got = precision.run_stats(s)
    assert s.err == 0
    s.close()
    assert ref["deposit_volume_m3"] > 0.0
    for k in ref:
        rel = abs(got[k] - ref[k]) / abs(ref[k])
        assert rel <= 1e-4, (k, got[k], ref[k], rel)
