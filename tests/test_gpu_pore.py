"""Pore pressure on the GPU (imex_pore.hip, through the C ABI) against the
oracle and the reference: the four cases with the physics of the
TEST_DAMBREAK decks (f_inhibit STATIC, DYNAMIC, OFF, STEP; mu(I) or Voellmy
friction; dynamic or constant permeability; gas loss) at 64 x 64 cells of
1 cm, device-driven for the fixture's steps from its q0.

Bars: bit for bit against the oracle's run with correctly rounded x**y, log,
log1p and exp (the rounding the device computes) -- dt sequence, active
cells, q, running maxima, error bits -- and within 1e-10 relative L-infinity
per variable over wet cells, with the same wet mask, against the reference
solver's own run (the C library's functions; tests/golden/make_pore.py)."""
import numpy as np
import pytest

from conftest import golden_path
import cases
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _fixture(name):
    return dict(np.load(golden_path(f"pore_{name}.npz"), allow_pickle=False))


def _static(g):
    return {k[len("static_"):]: v for k, v in g.items() if k.startswith("static_")}


@pytest.mark.parametrize("name", sorted(cases.PORE_CASES))
def test_pore_run_bit_exact_with_the_oracle(name):
    cfg, B, st, q0 = cases.PORE_CASES[name]()
    g = _fixture(name)
    steps = int(g["steps"])
    print(f"pore case {name}: {steps} device-driven steps", flush=True)
    s = Solver(cfg, _static(g), g["q0"], raise_on_error=False)
    s.err = 0
    dts = s.run_steps(steps)
    np.testing.assert_array_equal(dts, g["cr_dts"])
    q = s.q
    wet = q[..., 0] > 0.0
    if not np.array_equal(q, g["cr_q"]):
        d = np.argwhere(q != g["cr_q"])
        k, j, v = d[0]
        raise AssertionError(f"{len(d)} entries differ, first (k={k}, j={j}, var {v}): "
                             f"{q[k, j, v]!r} vs {g['cr_q'][k, j, v]!r}")
    f = s.fields()
    for key in ("hmax", "pdynmax", "mod_vel_max"):
        np.testing.assert_array_equal(f[key], g["cr_field_" + key], err_msg=key)
    assert s.err == int(g["cr_err"]) == 0
    # against the reference's own run: 1e-10 relative over wet cells
    qr = g["q"]
    assert np.array_equal(wet, qr[..., 0] > 0.0)
    scale = np.maximum(np.abs(qr[wet]).max(axis=0), 1e-300)
    rel = (np.abs(q[wet] - qr[wet]) / scale).max()
    assert rel <= 1e-10, rel
    assert np.max(np.abs(dts - g["dts"]) / g["dts"]) <= 1e-10
    s.close()


def test_pore_run_is_deterministic_and_steps_split_freely():
    """The same run in one call and in calls of 7 steps: the same bits."""
    cfg, B, st, q0 = cases.cp_64()
    g = _fixture("cp_64")
    a = Solver(cfg, _static(g), g["q0"], raise_on_error=False)
    da = a.run_steps(56)
    b = Solver(cfg, _static(g), g["q0"], raise_on_error=False)
    db = np.concatenate([b.run_steps(7) for _ in range(8)])
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(a.q, b.q)
    a.close()
    b.close()


def test_pore_host_driven_entry_points_refuse():
    """Runs with pore pressure step only through imex_run_steps: the
    host-driven entry points return an error instead of running kernels
    instantiated for the variables of runs without it."""
    from imex_sflow2d_dev_amd import native
    cfg, B, st, q0 = cases.cp_64()
    g = _fixture("cp_64")
    s = Solver(cfg, _static(g), g["q0"], raise_on_error=False)
    L = native.lib()
    assert L.imex_check_solve(s.h, 0) == -2
    assert L.imex_rk_step(s.h, native.ctypes.c_double(1e-4)) == -2
    s.close()


DAMBREAK_MODES = {"static": "STATIC", "dynamic": "DYNAMIC", "off": "OFF", "step": "STEP"}
# the correctly rounded run's distance from the reference's at T_END (see
# tests/test_oracle_pore.py): 1e-10; 5e-8 for the DYNAMIC deck, whose two
# roundings part at step 116 and end 2.1e-8 apart
DAMBREAK_TOL = {"static": 1e-10, "dynamic": 5e-8, "off": 1e-10, "step": 1e-10}


@pytest.mark.parametrize("key", sorted(DAMBREAK_MODES))
def test_dambreak_decks_on_the_gpu(key):
    """The reference's TEST_DAMBREAK decks (300 x 15 cells, their restart
    state, per-variable boundaries: h and h alpha_s one-sided slopes) run on
    the GPU to T_END = 0.3 s: bit for bit with the correctly rounded oracle,
    within 1e-10 of the reference's own run (5e-8 for the DYNAMIC deck)."""
    cfg, B, st, q0 = cases.dambreak_deck(DAMBREAK_MODES[key])
    g = dict(np.load(golden_path(f"dambreak_{key}.npz"), allow_pickle=False))
    steps = int(g["steps"])
    print(f"dambreak deck {key}: {steps} steps to t_end {cfg.t_end}", flush=True)
    s = Solver(cfg, st, q0, raise_on_error=False)
    s.err = 0
    dts = s.run_steps(steps + 3, t_end=cfg.t_end)
    np.testing.assert_array_equal(dts[:steps], g["cr_dts"])
    assert np.all(dts[steps:] == 0.0)  # the call stops at t_end
    np.testing.assert_array_equal(s.q, g["cr_q"])
    f = s.fields()
    for k in ("hmax", "pdynmax", "mod_vel_max"):
        np.testing.assert_array_equal(f[k], g["cr_field_" + k], err_msg=k)
    assert s.err == 0
    q, qr = s.q, g["q"]
    w = q[..., 0] > 0.0
    assert np.array_equal(w, qr[..., 0] > 0.0)
    scale = np.maximum(np.abs(qr[w]).max(axis=0), 1e-300)
    assert (np.abs(q[w] - qr[w]) / scale).max() <= DAMBREAK_TOL[key]
    s.close()
