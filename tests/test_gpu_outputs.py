"""Device-side outputs against the oracle: the runout reductions
(imex_runout_reduce), the probe values (imex_probes), the device-driven
steps' stop at t_end, and the runout and probe files of a run with outputs."""
import math
import os

import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init, outputs
from imex_sflow2d_dev_amd.solver import Solver
from oracle import oraclepy, outputs_ref

pytestmark = pytest.mark.gpu


def _solver(name, steps, precision=0):
    cfg, B = cases.CASES[name]()
    cfg.precision = precision
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    if steps:
        s.run_steps(steps)
    return cfg, B, st, q0, s


def _check_red(cfg, got, want):
    for k in ("n_hpos", "n_hpos_old", "n_new", "imax", "imax_x", "imax_y"):
        assert got[k] == want[k], k
    for k in ("max_h", "dist_max", "dist_x_max", "dist_y_max", "dist_at_imax_x", "dist_at_imax_y"):
        assert got[k] == want[k], k
    for k in ("sum_q1", "sum_xq1", "sum_yq1"):
        # double-double accumulation rounded once: within an ulp of the
        # correctly rounded sum, and within the reference's own (array-order)
        # rounding of its running sum
        assert abs(got[k] - want[k]) <= 2.0 * np.spacing(abs(want[k])), k
        assert abs(got[k] - want[k + "_naive"]) <= 1e-12 * abs(want[k]), k


@pytest.mark.parametrize("name, steps, precision", [("c1_64", 0, 0), ("c1_64", 300, 0),
                                                    ("c2_64", 200, 0), ("c2_64", 200, 1)])
def test_runout_reduce_matches_oracle(name, steps, precision):
    print(f"runout reduce {name} steps={steps} precision={precision}", flush=True)
    cfg, B, st, q0, s = _solver(name, steps, precision)
    try:
        o = oraclepy.Oracle(cfg)
        q = s.q
        qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
        x0, y0 = cfg.x0 + 0.37 * cfg.nx * cfg.dx, cfg.y0 + 0.61 * cfg.ny * cfg.dx
        old = np.zeros((cfg.ny, cfg.nx), bool)
        for thr in (0.0, 0.03):
            got = s.runout_reduce(x0, y0, thr)
            _check_red(cfg, got, outputs_ref.runout_red(cfg, q, qp, old, x0, y0, thr))
        s.set_topography(B)
        got = s.runout_reduce(x0, y0, 0.0, store=True, with_b=True)
        want = outputs_ref.runout_red(cfg, q, qp, old, x0, y0, 0.0, B=B)
        _check_red(cfg, got, want)
        assert got["ib"] == want["ib"] and got["b_max"] == want["b_max"]
        # the stored mask is the next call's hpos_old
        s.run_steps(40)
        q2 = s.q
        qp2 = outputs_ref.module_qp(cfg, q2, o.qc_to_qp)
        got = s.runout_reduce(x0, y0, 0.0, store=True)
        _check_red(cfg, got, outputs_ref.runout_red(cfg, q2, qp2, want["hpos"], x0, y0, 0.0))
        assert got["n_hpos"] > 0
    finally:
        s.close()


@pytest.mark.parametrize("name, steps", [("c1_64", 250), ("c2_64", 200)])
def test_probes_match_oracle(name, steps):
    print(f"probes {name} steps={steps}", flush=True)
    cfg, B, st, q0, s = _solver(name, steps)
    try:
        o = oraclepy.Oracle(cfg)
        q = s.q
        qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
        dep = s.fields()["deposit"]
        rng = np.random.default_rng(11)
        L = cfg.nx * cfg.dx
        x_comp = cfg.x0 + (np.arange(cfg.nx) + 0.5) * cfg.dx
        y_comp = cfg.y0 + (np.arange(cfg.ny) + 0.5) * cfg.dx
        wet = np.argwhere(q[:, :, 0] > 0)
        pts = [(x_comp[j] + rng.uniform(-5, 5), y_comp[k] + rng.uniform(-5, 5))
               for k, j in wet[rng.choice(len(wet), 60)]]
        pts += [(rng.uniform(-20, L + 20), rng.uniform(-20, L + 20)) for _ in range(100)]
        pts += [(x_comp[0], y_comp[0]), (x_comp[-1], y_comp[-1]), (L + 50.0, 50.0), (-3.0, 40.0)]
        xy = np.array(pts)
        got = s.probes(xy)
        want = outputs_ref.probes(cfg, q, qp, dep, xy)
        np.testing.assert_array_equal(got, want)
        assert np.count_nonzero(got[:, 0] > 1e-5) > 30
    finally:
        s.close()


def test_device_steps_stop_at_output_times():
    """run_steps with t_end at successive output times: each call ends on
    the output time, and the state, the dt sequence and the active-cell
    count equal the oracle's steps clipped at the same times."""
    cfg, B = cases.c2_64()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    try:
        s.run_steps(30)
        o.run(30)
        a0, _ = s.counters()
        targets = [s.t + 0.15, s.t + 0.15, s.t + 0.4, s.t + 0.41]
        dts_dev, dts_or, cells_or = [], [], 0
        for T in targets:
            print(f"segment to t = {T}", flush=True)
            d = s.run_steps(400, t_end=T)
            assert T <= s.t <= T + 1e-12
            dts_dev += list(d[d != 0.0])
            o.set_t_end(T)
            while o.t < T:
                dd, nc, err = o.run(1)
                dts_or += list(dd)
                cells_or += int(nc[0])
        assert dts_dev == dts_or
        np.testing.assert_array_equal(s.q, o.q())
        a1, _ = s.counters()
        assert a1 - a0 == cells_or
    finally:
        s.close()


def test_runout_and_probe_files_match_oracle(tmp_path):
    print("run with outputs c1_64", flush=True)
    cfg, B = cases.c1_64()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    dev_dir, ref_dir = tmp_path / "dev", tmp_path / "ref"
    dev_dir.mkdir()
    ref_dir.mkdir()
    o = oraclepy.Oracle(cfg)
    try:
        ro = outputs.RunoutOutput(s, "run", str(dev_dir), dt_runout=0.5, eps_stop=0.0)
        ref = outputs_ref.RunoutRef(cfg, o.qc_to_qp, str(ref_dir), "run", dt_runout=0.5)
        coords = [(200.0, 320.0), (260.0, 330.0), (5000.0, 10.0)]
        po = outputs.ProbesOutput(s, coords, "run", str(dev_dir), dt_probes=0.7)
        probe_rows = []
        seen = []

        def runout_hook(time, stop):
            q = s.q
            seen.append(time)
            ref(time, q, stop)
            return ro_call(time, stop)

        def probe_hook(time):
            q = s.q
            qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
            probe_rows.append((time, outputs_ref.probes(cfg, q, qp, s.fields()["deposit"], coords)))
            po_call(time)

        ro_call, po_call = ro.__call__, po.__call__
        steps, stop = outputs.run_with_outputs(
            s, 3.0, runout=_Wrap(ro, runout_hook), probes=_Wrap(po, probe_hook))
        assert steps > 50 and len(seen) == 7
        for name in ("run_runout.csv", "runout.txt"):
            assert (dev_dir / name).read_text() == (ref_dir / name).read_text(), name
        dev_mc = (dev_dir / "run_mass_center.csv").read_text().splitlines()
        ref_mc = (ref_dir / "run_mass_center.csv").read_text().splitlines()
        assert dev_mc[0] == ref_mc[0] and len(dev_mc) == len(ref_mc)
        for a, b in zip(dev_mc[1:], ref_mc[1:]):
            va = [float(x) for x in a.split(",")]
            vb = [float(x) for x in b.split(",")]
            assert max(abs(x - y) for x, y in zip(va, vb)) <= 0.0011
        # probe files: each line's values are the oracle's, written ES14.7E2
        for k in range(len(coords)):
            lines = (dev_dir / f"run_prb_{k + 1:04d}.csv").read_text().splitlines()
            assert len(lines) == 3 + len(probe_rows)
            for line, (t, rows) in zip(lines[3:], probe_rows):
                vals = [float(x) for x in line.split(",")]
                assert vals[0] == float(f"{t:.7E}")
                want = [float(f"{v:.7E}") for v in rows[k]]
                assert vals[1:] == want
    finally:
        s.close()


class _Wrap:
    """Routes calls to a hook while exposing the wrapped writer's state."""

    def __init__(self, inner, hook):
        self._inner, self._hook = inner, hook

    def __call__(self, *a):
        return self._hook(*a)

    def __getattr__(self, name):
        return getattr(self._inner, name)

    def __setattr__(self, name, value):
        if name in ("_inner", "_hook"):
            object.__setattr__(self, name, value)
        else:
            setattr(self._inner, name, value)
