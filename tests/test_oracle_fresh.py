"""The oracle's fresh mode (oracle_set_fresh): the states the reference keeps
from an earlier step for dry cells that left the active set (face states and
primitive states, src/solver_2d.f90 reconstruction :3095-3924 writes them for
active cells only) are read as a fresh evaluation gives them, which is what
the device computes.  Without such reads the two modes are the same run bit
for bit; with them (erosion runs, where cells of vanishing mass dry out and
come back) the state q stays the same and only the erosion, deposit and bed
of the cells next to the reads differ, by less than 1e-25 m here."""
import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init
from oracle import oraclepy


def _pair(name):
    cfg, B = cases.CASES[name]()
    st, q0 = init.setup(cfg, B)
    a = oraclepy.Oracle(cfg, st, q0)
    b = oraclepy.Oracle(cfg, st, q0)
    for o in (a, b):
        o.set_pow_mode(1)
    b.set_fresh(True)
    return a, b


@pytest.mark.parametrize("name, steps", [("c1_64", 300), ("c2_64", 200)])
def test_fresh_mode_is_the_same_run_without_stale_reads(name, steps):
    a, b = _pair(name)
    da, _, _ = a.run(steps)
    db, _, _ = b.run(steps)
    assert a.counters()[1] == 0
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(a.q(), b.q())
    np.testing.assert_array_equal(a.fields()["deposit"], b.fields()["deposit"])


def test_fresh_mode_with_stale_reads_changes_only_the_layers_of_their_cells():
    a, b = _pair("c2e_64")
    da, _, _ = a.run(300)
    db, _, _ = b.run(300)
    assert a.counters()[1] > 0 and b.counters()[1] == 0
    np.testing.assert_array_equal(da, db)
    np.testing.assert_array_equal(a.q(), b.q())
    stale = a.stale_cells()
    ea, eb = a.erosion()[0], b.erosion()[0]
    pa, pb = a.fields()["deposit"], b.fields()["deposit"]
    assert not np.array_equal(pa, pb)
    # outside the cells next to the reads, bit for bit
    assert np.array_equal(pa[:, ~stale], pb[:, ~stale])
    np.testing.assert_allclose(pa, pb, rtol=0, atol=1e-25)
    np.testing.assert_allclose(ea, eb, rtol=0, atol=1e-25)


def test_c2_trajectory_fixture_fresh_and_stale_readings_agree():
    """The two oracle runs of the committed 5000-step C2 fixture (the
    reference's stale-state reading and the fresh reading the device uses)
    are one run until the stale reads begin: the same dt, active cells and
    error bits up to step 4710, the same q at every checkpoint, the same
    deposit at every checkpoint before the first stale read, and the same
    NaN steps."""
    from conftest import golden_path
    g = np.load(golden_path("c2_trajectory.npz"), allow_pickle=False)
    n = 4710
    np.testing.assert_array_equal(g["fresh_dts"][:n], g["dts"][:n])
    np.testing.assert_array_equal(g["fresh_err"], g["err"])
    np.testing.assert_array_equal(g["fresh_ncells"][:n], g["ncells"][:n])
    np.testing.assert_array_equal(g["fresh_ck_q_sha"], g["ck_q_sha"])
    clean = g["ck_stale_reads"] == 0
    assert clean[:18].all() and not clean[18:].any()
    np.testing.assert_array_equal(g["fresh_ck_dep_sha"][clean], g["ck_dep_sha"][clean])
    np.testing.assert_array_equal(g["fresh_nan_step"], g["nan_step"])
