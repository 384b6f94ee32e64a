"""The first 120 steps of the 4096 x 4096 three-class workload (classes of
1e-2, 1e-3 and 1e-4 m, with settling) on the GPU against the oracle's run in
tests/golden/c3_prefix.npz (written on the CPU by
tests/golden/make_c3_prefix.py, correctly rounded x**y as on the device):
k_flux<3> and k_update<3, 7> at the precision form's production size.

Bar: bit for bit -- the dt, active-cell count and physics error bits of
every step, and the SHA-256 of q and of the deposit at steps 40, 80, 120."""
import hashlib
import os
import sys

import numpy as np
import pytest

from conftest import golden_path
from imex_sflow2d_dev_amd import init
from imex_sflow2d_dev_amd.solver import Solver

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_c3_prefix  # noqa: E402  (the case definition only)

pytestmark = pytest.mark.gpu


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def test_c3_prefix_matches_oracle():
    g = np.load(golden_path("c3_prefix.npz"), allow_pickle=False)
    cfg, B = make_c3_prefix.case()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0, raise_on_error=False)
    ck = {int(c): i for i, c in enumerate(g["checkpoints"])}
    a_prev, _ = s.counters()
    for n in range(int(g["steps"])):
        s.err = 0
        dt = s.run_steps(1)[0]
        a, _ = s.counters()
        assert dt == g["dts"][n] and s.err == int(g["err"][n]), f"step {n + 1}"
        assert a - a_prev == int(g["ncells"][n]), f"step {n + 1}: active cells"
        a_prev = a
        if n + 1 in ck:
            i = ck[n + 1]
            q = s.q
            dep = np.ascontiguousarray(np.transpose(s.fields()["deposit"], (1, 2, 0)))
            assert _sha(q) == g["ck_q_sha"][i], f"q differs at step {n + 1}"
            assert _sha(dep) == g["ck_dep_sha"][i], f"deposit differs at step {n + 1}"
            assert int(np.count_nonzero(q[:, :, 0] > 0.0)) == int(g["ck_n_wet"][i])
    s.close()
