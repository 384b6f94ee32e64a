"""The driver's stdout diagnostics (src/IMEX_SfloW2D.f90:507-544) from
imex_post_step: mass dx*dy*SUM(q(1,:,:)), volume dx*dy*SUM(qp(1,:,:)) over
wet cells, area dx*dy*COUNT(q(1,:,:) > 1e-7) and solid mass
dx*dy*SUM(q(5:4+n_solid,:,:)), summed on the device with Neumaier
compensation.

Bar: each equals dx*dy times the correctly rounded sum (math.fsum) of the
oracle's state after the same steps to within 4 ulp (the device state
equals the oracle's bit for bit; the compensated sum is within about one
ulp of the correctly rounded one, and the product by dx*dy rounds once
more).  The per-class solid budget SUM(q_s) + rho_s*SUM(deposit_s) of a
collapse with settling is constant to round-off on the device while no cell
lifts off (the mixture's reduced gravity going non-positive empties a cell,
src/solver_2d.f90:2573-2593), and for the coarse class, which deposits
before the flow gets buoyant, over the whole run."""
import math

import numpy as np
import pytest

import cases
from oracle import oraclepy
from imex_sflow2d_dev_amd import init
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _fsum(a):
    return math.fsum(np.asarray(a, np.float64).ravel().tolist())


def _close(dev, ref, what, ulps=4):
    tol = ulps * np.spacing(abs(ref)) if ref != 0.0 else 0.0
    assert abs(dev - ref) <= tol, f"{what}: {dev!r} vs correctly rounded {ref!r}"


@pytest.mark.parametrize("name, steps", [("c1_64", 300), ("c2_64", 300), ("c1d_96", 200)])
def test_stats_equal_correctly_rounded_sums(name, steps):
    cfg, B = cases.CASES[name]()
    st, q0 = init.setup(cfg, B)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    s = Solver(cfg, st, q0)
    s.run_steps(steps)
    o.run(steps)
    q = o.q()
    assert np.array_equal(s.q, q), "device state differs from the oracle's"
    stats = s.stats()
    a = cfg.dx * cfg.dx
    wet = q[:, :, 0] > 0.0
    qp, _ = o.qc_to_qp(q[wet])
    _close(stats.mass, a * _fsum(q[:, :, 0]), "mass")
    _close(stats.volume, a * _fsum(qp[:, 0]), "volume")
    _close(stats.area, a * float(np.count_nonzero(q[:, :, 0] > 1.0e-7)), "area", ulps=0)
    _close(stats.solid_mass, a * _fsum(q[:, :, 4:4 + cfg.n_solid]), "solid mass")
    s.close()


def _budget(cfg, q, dep):
    return [_fsum(q[:, :, 4 + i]) + cfg.rho_s[i] * _fsum(dep[i]) for i in range(cfg.n_solid)]


def test_solid_budget_per_class_closes():
    cfg, B = cases.c1d_96()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    b0 = _budget(cfg, s.q, s.fields()["deposit"])
    s.run_steps(50)                      # before any cell lifts off: both classes
    b = _budget(cfg, s.q, s.fields()["deposit"])
    for i in range(cfg.n_solid):
        assert abs(b[i] - b0[i]) <= 1e-14 * b0[i], (i, b[i], b0[i])
    s.run_steps(250)                     # the coarse class over the whole run
    b = _budget(cfg, s.q, s.fields()["deposit"])
    assert abs(b[0] - b0[0]) <= 1e-14 * b0[0], (b[0], b0[0])
    assert np.count_nonzero(s.fields()["deposit"][0]) > 0
    s.close()
