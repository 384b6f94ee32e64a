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
