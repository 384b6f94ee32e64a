"""Size-independent properties of the HIP path at the full 2048 x 2048
production size, where the oracle is too slow to follow a long run."""
import numpy as np
import pytest

from conftest import golden_path
from imex_sflow2d_dev_amd import init, workloads
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c2():
    cfg, B = workloads.config2()
    st, q0 = init.setup(cfg, B)
    return cfg, st, q0


def active_set(cfg, st, q):
    """check_solve restated with numpy: wet interior cells, the source ring
    and the boundary frame, dilated twice by the 4-neighbourhood inside the
    interior, minus the source disc (src/solver_2d.f90:667-858)."""
    ny, nx = cfg.ny, cfg.nx
    m = np.zeros((ny, nx), bool)
    m[[0, -1], :] = True
    m[:, [0, -1]] = True
    inner = (q[1:-1, 1:-1, 0] > 0.0) | (st["source_cell"][1:-1, 1:-1] == 2)
    m[1:-1, 1:-1] |= inner
    for _ in range(2):
        n = m.copy()
        n[1:-1, 1:-1] |= m[1:-1, :-2] | m[1:-1, 2:] | m[:-2, 1:-1] | m[2:, 1:-1]
        m = n
    m[st["source_cell"] == 1] = False
    return m


def test_deterministic(c2):
    cfg, st, q0 = c2
    a = Solver(cfg, st, q0)
    da = a.run_steps(400)
    qa = a.q
    a.close()
    b = Solver(cfg, st, q0)
    db = b.run_steps(400)
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(b.q, qa)
    b.close()


def test_active_set_rule_and_finite_state(c2):
    """Also: the physics error bits of each 300-step block are those of the
    oracle's trajectory (tests/golden/c2_trajectory.npz): a NaN in the
    x-direction solid-flux cap (src/solver_2d.f90:2903-2912, a sum of +Inf
    and -Inf when the KT mass flux of a face is exactly zero) at steps 582
    and 623 of this run, nothing else; the cell is emptied by the lift-off
    rule of the same step and the state stays finite."""
    g = np.load(golden_path("c2_trajectory.npz"), allow_pickle=False)
    cfg, st, q0 = c2
    s = Solver(cfg, st, q0, raise_on_error=False)
    for b in range(3):
        s.err = 0
        s.run_steps(300)
        assert s.err == int(np.bitwise_or.reduce(g["err"][300 * b:300 * (b + 1)])), b
        q = s.q
        assert np.isfinite(q).all()
        s.check_solve(False)              # the set for the current state
        np.testing.assert_array_equal(s.mask().astype(bool), active_set(cfg, st, q))
        wet = q[:, :, 0] > 0.0
        assert (q[wet, 0] > 0.0).all() and (q[~wet] == 0.0).all()
        assert (q[:, :, 4:] >= 0.0).all()
    s.close()


def test_dt_output_after_a_timed_run_without_it():
    """imex_run_steps_timed with no dt output between two imex_run_steps
    calls that ask for it: the later call still returns the dt sequence, and
    the run equals one driven by imex_run_steps alone."""
    import ctypes
    from imex_sflow2d_dev_amd import native
    import cases
    cfg, B = cases.CASES["c2_96"]()
    st, q0 = init.setup(cfg, B)
    a = Solver(cfg, st, q0)
    b = Solver(cfg, st, q0)
    a.run_steps(40)
    b.run_steps(40)
    k = np.zeros(3)
    c = np.zeros(1, np.int64)
    ns = ctypes.c_int()
    rc = a.L.imex_run_steps_timed(a.h, 20, ctypes.c_double(cfg.t_end), None, ctypes.c_int(3),
                                  native._dp(k), ctypes.byref(ns),
                                  c.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
    assert rc == 0
    assert (k > 0.0).all() and c[0] > 0 and ns.value == 7
    b.run_steps(20)
    da = a.run_steps(30)
    db = b.run_steps(30)
    assert np.array_equal(da, db) and (da > 0.0).all()
    assert np.array_equal(a.q, b.q)
    a.close()
    b.close()
