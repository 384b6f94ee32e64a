"""Erosion of an erodible layer on the device (the erosion_coeff > 0 branch
of eval_mass_exchange_terms, src/constitutive_2d.f90:4731-4758, and the
erodible / erosion updates of update_erosion_deposition_cell,
src/solver_2d.f90:2478-2491) against the oracle, which
tests/test_oracle_vs_ref.py pins to the reference for these cases."""
import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init
from imex_sflow2d_dev_amd.solver import KernelProbe, Solver
from oracle import oraclepy

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name, steps, driver, deposit_flag, topo", [
    ("c1e_48", 300, "device", False, False), ("c1e_48", 120, "host", False, False),
    ("c2e_64", 300, "device", False, False), ("c2e_64", 200, "device", True, False),
    ("c1e_48", 300, "device", False, True), ("c1e_48", 120, "host", False, True),
    ("c2e_64", 300, "device", True, True)])
def test_erosion_run_matches_oracle(name, steps, driver, deposit_flag, topo):
    print(f"erosion {name} steps={steps} {driver} erodible_deposit={deposit_flag} "
          f"topo_change={topo}", flush=True)
    cfg, B = cases.CASES[name]()
    cfg.erodible_deposit_flag = deposit_flag
    cfg.topo_change_flag = topo
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    # With erosion, cells of vanishing mass dry out and come back often, and
    # the reference then reads face states left from an earlier step; the
    # device evaluates them fresh, as the oracle does in its fresh mode
    # (tests/test_oracle_fresh.py bounds the difference of the two readings).
    o.set_fresh(True)
    try:
        if driver == "device":
            d = s.run_steps(steps)
        else:
            d = np.array([s.step() for _ in range(steps)])
        do, _, err = o.run(steps)
        np.testing.assert_array_equal(d, do)
        np.testing.assert_array_equal(s.q, o.q())
        ers_d, erd_d = s.erosion()
        ers_o, erd_o = o.erosion()
        np.testing.assert_array_equal(erd_d, erd_o)
        np.testing.assert_array_equal(ers_d, ers_o)
        np.testing.assert_array_equal(s.fields()["deposit"], o.fields()["deposit"])
        assert ers_o.max() > 1e-4
        if topo:
            np.testing.assert_array_equal(s.topography(), o.topography())
            assert not np.array_equal(o.topography(), B)
    finally:
        s.close()


@pytest.mark.parametrize("name", ["c1e_48", "c2e_64"])
def test_mass_exchange_with_erosion_matches_oracle(name):
    cfg, B = cases.CASES[name]()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    try:
        s.run_steps(150)
        q = s.q.reshape(-1, cfg.n_vars)
        wet = q[q[:, 0] > 0.0]
        o = oraclepy.Oracle(cfg)
        o.set_pow_mode(1)          # x**y correctly rounded, as on the device
        qp, _ = o.qc_to_qp(wet)
        rng = np.random.default_rng(9)
        # layers from exhausted to thick, and zero
        erod = rng.choice([0.0, 1e-9, 1e-4, 0.05, 0.4], size=(len(wet), cfg.n_solid))
        kp = KernelProbe(cfg)
        for dt in (0.03, 2.0):
            got = kp.mass_exchange_ers(qp, erod, dt)
            want = o.mass_exchange_ers(qp, erod, dt)
            for a, b in zip(got, want):
                np.testing.assert_array_equal(a, b)
            assert (want[2] > 0).any() and (want[2] == erod / dt).any()
        kp.close()
    finally:
        s.close()
