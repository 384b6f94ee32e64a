"""imex_post_step_noalloc, the allocation-free variant of imex_post_step
(the driver epilogue and the stdout diagnostics with the partial sums in a
static device array), through the C ABI against imex_post_step on the same
state: two calls in a row on two runs, the second on a smaller grid, each
equal bit for bit to imex_post_step's statistics."""
import ctypes

import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init, native
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _stats(L, h, fn, t):
    st = native.ImexStepStats()
    rc = getattr(L, fn)(h, ctypes.c_double(t), ctypes.byref(st))
    assert rc >= 0, (fn, rc)
    return st


def _fields(st):
    return (st.t, st.dt, st.mass, st.volume, st.area, st.solid_mass, st.active_cells,
            st.dried_cells)


@pytest.mark.parametrize("order", [("c2_96", "c1_64"), ("c1d_96", "c1_48")])
def test_noalloc_post_step_equals_post_step(order):
    L = native.lib()
    for name in order:                  # the second run has the smaller grid
        cfg, B = cases.CASES[name]()
        st, q0 = init.setup(cfg, B)
        s = Solver(cfg, st, q0)
        s.run_steps(120)
        a = _stats(L, s.h, "imex_post_step_noalloc", s.t)
        b = _stats(L, s.h, "imex_post_step", s.t)
        c = _stats(L, s.h, "imex_post_step_noalloc", s.t)
        fa, fb, fc = _fields(a), _fields(b), _fields(c)
        assert fa == fb == fc, (name, fa, fb)
        assert np.float64(a.mass).tobytes() == np.float64(b.mass).tobytes()
        assert a.mass > 0.0 and a.area > 0.0
        s.close()


def test_noalloc_post_step_without_stats():
    L = native.lib()
    cfg, B = cases.c1_64()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    s.run_steps(20)
    assert L.imex_post_step_noalloc(s.h, ctypes.c_double(s.t), None) == 0
    s.close()
