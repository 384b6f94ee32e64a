"""The whole 5000-step 2048 x 2048 radial-source run (two solid classes,
settling and air entrainment) on the GPU against the oracle's trajectory in
tests/golden/c2_trajectory.npz (written on the CPU by
tests/golden/make_c2_trajectory.py: the C oracle with correctly rounded x**y,
the rounding the HIP path computes).

The fixture holds two oracle runs: the reference's own reading, in which a
dry cell outside the active set keeps face and stage states from the step it
was last active (the stale-state quirk; the reads of such states that differ
from a fresh evaluation begin at step 4379, as the flow lifts off), and the
run that evaluates those states fresh, as the device does (fresh_*).  The two
are the same run up to step 4710 and in q at every checkpoint.

Bar: bit for bit with the fresh run at every step -- the dt sequence, the
active-cell count and the physics error bits of each step (the steps whose
assembled state held a NaN, where the reference prints the cell and pauses,
src/solver_2d.f90:1382-1397, are exactly the oracle's) -- and the SHA-256 of
q and of the deposit at the checkpoint steps; and with the reference's
reading, q at every checkpoint and the deposit outside the cells next to a
stale read."""
import hashlib

import numpy as np
import pytest

from conftest import golden_path
from imex_sflow2d_dev_amd import init, workloads
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu

# the fresh and the stale reading of the fixture take the same dt to here
SAME_STEPS = 4710


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
        if dts[n] != g["fresh_dts"][n] or err[n] != g["fresh_err"][n]:
            pytest.fail(f"step {n + 1}: dt {dts[n]!r} vs {g['fresh_dts'][n]!r}, error bits "
                        f"{err[n]} vs {g['fresh_err'][n]}")
        if n + 1 in ck:
            i = ck[n + 1]
            q = s.q
            dep = np.ascontiguousarray(np.transpose(s.fields()["deposit"], (1, 2, 0)))
            assert _sha(q) == g["fresh_ck_q_sha"][i], f"q differs at step {n + 1}"
            assert _sha(dep) == g["fresh_ck_dep_sha"][i], f"deposit differs at step {n + 1}"
            assert int(np.count_nonzero(q[:, :, 0] > 0.0)) == int(g["fresh_ck_n_wet"][i])
            # the reference's reading: the same q, and, while the two
            # readings take the same steps (to step 4710), the same deposit
            # outside the cells next to a stale read
            assert _sha(q) == g["ck_q_sha"][i], f"q differs from the stale reading at {n + 1}"
            if n + 1 > SAME_STEPS:
                continue
            a = int(np.sum(g["ck_stale_cells_n"][:i]))
            cells = g["ck_stale_cells"][a:a + int(g["ck_stale_cells_n"][i])]
            stale = np.zeros(q.shape[0] * q.shape[1], bool)
            stale[cells] = True
            stale = stale.reshape(q.shape[:2])
            masked = np.where(stale[:, :, None], 0.0, dep)
            assert _sha(masked) == g["ck_dep_fresh_sha"][i], \
                f"deposit outside the stale-read cells differs at step {n + 1}"
    np.testing.assert_array_equal(ncells, g["fresh_ncells"])
    # the NaN events are the oracle's, step for step
    np.testing.assert_array_equal(np.nonzero(err & 1)[0] + 1, g["fresh_nan_step"])
    s.close()
