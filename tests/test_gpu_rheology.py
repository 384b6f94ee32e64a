"""The mu(I) rheologies on the device: model 11 (mu(I) friction in the
semi-implicit terms, src/constitutive_2d.f90:4384-4458, no implicit term)
and model 12 (the same friction plus the Voellmy turbulent term,
:4461-4537 and :3660-3683), over multi-step runs against the oracle, which
tests/test_oracle_vs_ref.py pins bit for bit to the reference for these
cases."""
import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init
from imex_sflow2d_dev_amd.solver import Solver
from oracle import oraclepy

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name, steps, driver", [
    ("c1m_64", 400, "device"), ("c1m_64", 150, "host"),
    ("c2m_64", 400, "device"), ("c2m_64", 150, "host")])
def test_mu_i_run_matches_oracle(name, steps, driver):
    print(f"mu(I) {name} steps={steps} {driver}", flush=True)
    cfg, B = cases.CASES[name]()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    try:
        if driver == "device":
            d = s.run_steps(steps)
        else:
            d = np.array([s.step() for _ in range(steps)])
        do, _, err = o.run(steps)
        np.testing.assert_array_equal(d, do)
        np.testing.assert_array_equal(s.q, o.q())
        fd, fo = s.fields(), o.fields()
        for k in ("hmax", "pdynmax", "mod_vel_max", "deposit"):
            np.testing.assert_array_equal(fd[k], fo[k])
        # the flow moved: friction of the mu(I) law, not a dry or frozen grid
        assert fo["mod_vel_max"].max() > 1.0
    finally:
        s.close()


def test_mu_i_differs_from_voellmy():
    """Switching the friction law changes the run (the model field reaches
    the kernels): the same case under model 1 and model 11 diverges."""
    cfg, B = cases.CASES["c1m_64"]()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    cfg1, _ = cases.CASES["c1m_64"]()
    cfg1.rheology_model = 1
    s1 = Solver(cfg1, st, q0)
    try:
        s.run_steps(100)
        s1.run_steps(100)
        assert not np.array_equal(s.q, s1.q)
    finally:
        s.close()
        s1.close()


@pytest.mark.parametrize("name, steps, driver", [
    ("c1s_64", 300, "device"), ("c1s_64", 120, "host"), ("c2s_64", 300, "device")])
def test_stochastic_friction_run_matches_oracle(name, steps, driver):
    """The stochastic friction: each cell's xi + Z limited to [1, 1e5]
    (src/constitutive_2d.f90:3663-3673) with the Z given by imex_set_noise;
    the oracle takes the same Z and tests/test_oracle_vs_ref.py pins it to
    the reference."""
    print(f"stochastic {name} steps={steps} {driver}", flush=True)
    cfg, B = cases.CASES[name]()
    Z = cases.stochastic_noise(cfg)
    st, q0 = init.setup(cfg, B)
    st["Z"] = Z
    s = Solver(cfg, st, q0)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    try:
        if driver == "device":
            d = s.run_steps(steps)
        else:
            d = np.array([s.step() for _ in range(steps)])
        do, _, err = o.run(steps)
        np.testing.assert_array_equal(d, do)
        np.testing.assert_array_equal(s.q, o.q())
    finally:
        s.close()


def test_set_noise_needs_the_stochastic_flag():
    cfg, B = cases.CASES["c1_64"]()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    try:
        with pytest.raises(Exception, match="stochastic"):
            s.set_noise(np.zeros((cfg.ny, cfg.nx)))
    finally:
        s.close()
