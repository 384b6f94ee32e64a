"""The transported noise on the GPU (imex_pore.hip's general path through
the C ABI) against the reference's own runs in
tests/golden/transport_<case>.npz (tests/golden/make_transport.py): a
one-class collapse with the stochastic Voellmy friction, the same with pore
pressure and gas loss (the noise's mass-exchange source), and a two-class
collapse (the 2 x 2 Newton system with two solid rows and the noise row).

Bar: bit for bit -- dt sequence, q, running maxima, error bits.  The
oracle's correctly rounded run (the rounding the device computes) equals the
reference's on these cases (tests/test_oracle_transport.py), so the device
is held to the reference itself."""
import ctypes

import numpy as np
import pytest

from conftest import golden_path
import cases
from imex_sflow2d_dev_amd import native
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _fixture(name):
    return dict(np.load(golden_path(f"transport_{name}.npz"), allow_pickle=False))


def _static(g):
    return {k[len("static_"):]: v for k, v in g.items() if k.startswith("static_")}


@pytest.mark.parametrize("name", sorted(cases.TRANSPORT_CASES))
def test_transport_run_bit_exact_with_the_reference(name):
    cfg, B, st, q0, Z = cases.TRANSPORT_CASES[name]()
    g = _fixture(name)
    steps = int(g["steps"])
    print(f"transport case {name}: {steps} device-driven steps", flush=True)
    s = Solver(cfg, _static(g), g["q0"], raise_on_error=False)
    s.set_noise(g["Z"])
    s.err = 0
    dts = s.run_steps(steps)
    np.testing.assert_array_equal(dts, g["cr_dts"])
    np.testing.assert_array_equal(dts, g["dts"])
    q = s.q
    qr = g.get("cr_q", g["q"])
    if not np.array_equal(q, qr):
        d = np.argwhere(q != qr)
        k, j, v = d[0]
        raise AssertionError(f"{len(d)} entries differ, first (k={k}, j={j}, var {v}): "
                             f"{q[k, j, v]!r} vs {qr[k, j, v]!r}")
    f = s.fields()
    for key in ("hmax", "pdynmax", "mod_vel_max"):
        np.testing.assert_array_equal(f[key], g.get("cr_field_" + key, g["field_" + key]),
                                      err_msg=key)
    assert s.err == int(g["cr_err"]) == 0
    s.close()


def test_transport_runs_refuse_what_the_general_path_does_not_carry():
    """imex_create rejects the transported noise without the stochastic
    friction or with settling; the host-driven entry points refuse such
    runs."""
    cfg, B, st, q0, Z = cases.c1t_64()
    L = native.lib()
    h = ctypes.c_void_p()
    c = native.make_config(cfg)
    c.stochastic_flag = 0
    assert L.imex_create(ctypes.byref(c), ctypes.byref(h)) == -3
    c = native.make_config(cfg)
    c.settling_flag = 1
    assert L.imex_create(ctypes.byref(c), ctypes.byref(h)) == -3
    s = Solver(cfg, st, q0, raise_on_error=False)
    assert L.imex_check_solve(s.h, 0) == -2
    assert L.imex_rk_step(s.h, ctypes.c_double(1e-4)) == -2
    s.close()


def test_transport_run_splits_freely_and_constant_noise_stays_proportional():
    """The two-class case in one call of 300 steps and in calls of 50: the
    same bits.  And with Z one constant c everywhere, the device's q1 * Z
    stays c q1 (advection and mass exchange alike) within 1e-12 of the
    largest c q1 after 300 steps, dry cells 0."""
    cfg, B, st, q0, Z = cases.c2t_48()
    a = Solver(cfg, st, q0, raise_on_error=False)
    a.set_noise(Z)
    da = a.run_steps(300)
    b = Solver(cfg, st, q0, raise_on_error=False)
    b.set_noise(Z)
    db = np.concatenate([b.run_steps(50) for _ in range(6)])
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(a.q, b.q)
    a.close()
    b.close()
    c = 37.25
    ist = 4 + cfg.n_solid
    qc = q0.copy()
    qc[..., ist] = qc[..., 0] * c
    s = Solver(cfg, st, qc, raise_on_error=False)
    s.set_noise(np.full_like(Z, c))
    s.err = 0
    s.run_steps(300)
    q = s.q
    assert s.err == 0
    scale = np.max(np.abs(c * q[..., 0]))
    assert np.max(np.abs(q[..., ist] - c * q[..., 0])) <= 1e-12 * scale
    assert np.all(q[q[..., 0] == 0.0][:, ist] == 0.0)
    assert not np.array_equal(q[..., 0], qc[..., 0])
    s.close()
