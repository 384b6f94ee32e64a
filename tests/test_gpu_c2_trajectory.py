"""The whole 5000-step 2048 x 2048 radial-source run (two solid classes,
settling and air entrainment) on the GPU against the oracle's trajectory in
tests/golden/c2_trajectory.npz (written on the CPU by
tests/golden/make_c2_trajectory.py: the C oracle with correctly rounded x**y,
the rounding the HIP path computes).

Bar: bit for bit at every step -- the dt sequence, the active-cell count and
the physics error bits of each step (the steps whose assembled state held a
NaN, where the reference prints the cell and pauses,
src/solver_2d.f90:1382-1397, are exactly the oracle's) -- and the SHA-256 of
q and of the deposit at the checkpoint steps."""
import hashlib

import numpy as np
import pytest

from conftest import golden_path
from imex_sflow2d_dev_amd import init, workloads
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c2_whole_run_matches_oracle_trajectory():
    g = np.load(golden_path("c2_trajectory.npz"), allow_pickle=False)
    steps = int(g["steps"])
    ck = {int(s): i for i, s in enumerate(g["checkpoints"])}
    cfg, B = workloads.config2()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0, raise_on_error=False)
    a_prev, _ = s.counters()
    dts = np.zeros(steps)
    err = np.zeros(steps, np.uint8)
    ncells = np.zeros(steps, np.int64)
    for n in range(steps):
        s.err = 0
        dts[n] = s.run_steps(1)[0]
        err[n] = s.err
        a, _ = s.counters()
        ncells[n] = a - a_prev
        a_prev = a
        if dts[n] != g["dts"][n] or err[n] != g["err"][n]:
            pytest.fail(f"step {n + 1}: dt {dts[n]!r} vs {g['dts'][n]!r}, error bits "
                        f"{err[n]} vs {g['err'][n]}")
        if n + 1 in ck:
            i = ck[n + 1]
            q = s.q
            dep = np.ascontiguousarray(np.transpose(s.fields()["deposit"], (1, 2, 0)))
            assert _sha(q) == g["ck_q_sha"][i], f"q differs at step {n + 1}"
            assert _sha(dep) == g["ck_dep_sha"][i], f"deposit differs at step {n + 1}"
            assert int(np.count_nonzero(q[:, :, 0] > 0.0)) == int(g["ck_n_wet"][i])
    np.testing.assert_array_equal(ncells, g["ncells"])
    # the NaN events are the oracle's, step for step
    np.testing.assert_array_equal(np.nonzero(err & 1)[0] + 1, g["nan_step"])
    s.close()
