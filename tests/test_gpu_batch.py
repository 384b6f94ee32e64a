"""Batches of independent runs (imex_batch_*): every member's result is the
same, bit for bit, as the same run alone through the single-run device
driver -- members of different grid sizes and parameters, one that joins
the batch after it started, and two-class members with settling and the
radial source."""
import numpy as np
import pytest

from imex_sflow2d_dev_amd import init, workloads
from imex_sflow2d_dev_amd.ensemble import Batch, run_alone
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _bits(a, b, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{what}: {int((~same).sum())} mismatches"


def _same_run(s, r, what):
    assert s.t == r.t, what
    _bits(s.q, r.q, what + " q")
    fs, fr = s.fields(), r.fields()
    for k in ("hmax", "pdynmax", "mod_vel_max", "deposit"):
        _bits(fs[k], fr[k], f"{what} {k}")
    assert np.array_equal(s.mask(), r.mask()), what + " active set"


def _members(n, t_end, sides=(40, 90)):
    out = []
    for m in range(n):
        cfg, B = workloads.config4_member(m, min_side=sides[0], max_side=sides[1], t_end=t_end)
        st, q0 = init.setup(cfg, B)
        out.append((cfg, st, q0))
    return out


def test_members_match_runs_alone():
    mem = _members(6, t_end=4.0)
    alone = []
    for cfg, st, q0 in mem:
        s = Solver(cfg, st, q0)
        run_alone(s, cfg.t_end)
        alone.append(s)
    b = Batch(n_solid=1, blocks_per_member=8)
    inb = []
    for cfg, st, q0 in mem:
        s = Solver(cfg, st, q0)
        b.add(s, cfg.t_end)
        inb.append(s)
    steps = b.run_to_end(chunk=25)
    done, nst, t = b.status()
    assert done.all() and steps > 0 and (nst > 0).all()
    for s in list(inb):
        b.remove(s)
    for i, (s, r) in enumerate(zip(inb, alone)):
        assert s.t >= mem[i][0].t_end
        _same_run(s, r, f"member {i}")
    b.close()


def test_member_joining_later_matches_alone():
    mem = _members(3, t_end=3.0)
    b = Batch(n_solid=1, blocks_per_member=16)
    first = [Solver(c, st, q0) for c, st, q0 in mem[:2]]
    for s, (c, _, _) in zip(first, mem[:2]):
        b.add(s, c.t_end)
    b.run(40)
    late = Solver(*mem[2])
    b.add(late, mem[2][0].t_end)
    b.run_to_end(chunk=20)
    b.remove(late)
    ref = Solver(*mem[2])
    run_alone(ref, mem[2][0].t_end)
    _same_run(late, ref, "late member")
    b.close()


def test_two_class_radial_source_members():
    runs = []
    for nx, dx in ((64, 25.0), (80, 20.0)):
        cfg, B = workloads.small_config2(nx=nx, ny=nx, dx=dx)
        cfg.t_end = 6.0
        st, q0 = init.setup(cfg, B)
        runs.append((cfg, st, q0))
    b = Batch(n_solid=2, blocks_per_member=8)
    inb = [Solver(*r) for r in runs]
    for s, r in zip(inb, runs):
        b.add(s, r[0].t_end)
    b.run_to_end(chunk=30)
    for s, r in zip(list(inb), runs):
        b.remove(s)
        ref = Solver(*r)
        run_alone(ref, r[0].t_end)
        _same_run(s, ref, f"two-class {r[0].nx}")
    b.close()


def test_served_results_equal_runs_alone():
    """The serving loop (arrivals, admission under a small memory budget so
    that requests wait and join as others leave): every served result equals
    the same request run alone."""
    from imex_sflow2d_dev_amd import serving
    reqs = serving.make_requests(rate=40.0, duration=0.3, seed=9, t_end=(2.0, 5.0), side=(40, 90))
    assert len(reqs) >= 4
    budget = 2.5 * max(r.footprint for r in reqs)
    serving.serve(reqs, budget_bytes=budget, chunk=7, keep=True)
    for r in reqs:
        s = Solver(r.cfg, r.static, r.q0)
        run_alone(s, r.t_end)
        assert s.t == r.t, r.idx
        _bits(s.q, r.q, f"request {r.idx}")
        s.close()


def test_served_two_class_requests_equal_runs_alone():
    """Requests with two solid classes and settling (a batch of n_solid = 2):
    every served result equals the same request run alone."""
    from imex_sflow2d_dev_amd import serving
    reqs = serving.make_requests(rate=40.0, duration=0.25, seed=13, t_end=(2.0, 4.0),
                                 side=(40, 80), n_solid=2)
    assert len(reqs) >= 3 and all(r.cfg.n_solid == 2 and r.cfg.settling_flag for r in reqs)
    budget = 2.5 * max(r.footprint for r in reqs)
    serving.serve(reqs, budget_bytes=budget, chunk=6, keep=True)
    for r in reqs:
        s = Solver(r.cfg, r.static, r.q0)
        run_alone(s, r.t_end)
        assert s.t == r.t, r.idx
        _bits(s.q, r.q, f"request {r.idx}")
        s.close()


def test_serving_admits_an_over_budget_request_alone():
    """Admission accounts each run at its handle's real device footprint
    (imex_get_footprint); a request larger than the whole budget runs alone
    once nothing else runs, and the queue drains."""
    from imex_sflow2d_dev_amd import serving
    reqs = serving.make_requests(rate=40.0, duration=0.2, seed=11, t_end=(1.0, 2.0),
                                 side=(40, 70))
    assert len(reqs) >= 2
    s = Solver(reqs[0].cfg, reqs[0].static, reqs[0].q0)
    fp = s.footprint()
    s.close()
    N = reqs[0].cfg.nx * reqs[0].cfg.ny
    assert fp >= 8 * (2 * reqs[0].cfg.n_vars + 4) * N
    budget = 0.5 * min(r.footprint for r in reqs)
    serving.serve(reqs, budget_bytes=budget, chunk=5, keep=True)
    assert all(r.over_budget and r.finish is not None for r in reqs)
    assert reqs[0].footprint == fp
    # one at a time: each admission follows the previous result
    order = sorted(reqs, key=lambda r: r.admit)
    for a, b in zip(order, order[1:]):
        assert b.admit >= a.finish
    for r in reqs:
        s = Solver(r.cfg, r.static, r.q0)
        run_alone(s, r.t_end)
        _bits(s.q, r.q, f"request {r.idx}")
        s.close()


def test_members_with_different_friction_laws():
    """Voellmy, mu(I) (model 11) and stochastic-friction members in one
    batch: each member's friction law and noise travel with it."""
    import cases
    specs = []
    for name in ("c1_64", "c1m_64", "c1s_64"):
        cfg, B = cases.CASES[name]()
        cfg.t_end = 3.0
        st, q0 = init.setup(cfg, B)
        if cfg.stochastic_flag:
            st["Z"] = cases.stochastic_noise(cfg)
        specs.append((name, cfg, st, q0))
    alone = []
    for name, cfg, st, q0 in specs:
        s = Solver(cfg, st, q0)
        run_alone(s, cfg.t_end)
        alone.append(s)
    b = Batch(n_solid=1, blocks_per_member=8)
    inb = []
    for name, cfg, st, q0 in specs:
        s = Solver(cfg, st, q0)
        b.add(s, cfg.t_end)
        inb.append(s)
    b.run_to_end(chunk=25)
    for s in list(inb):
        b.remove(s)
    for (name, *_), s, r in zip(specs, inb, alone):
        _same_run(s, r, name)
    assert not np.array_equal(inb[0].q, inb[1].q) and not np.array_equal(inb[0].q, inb[2].q)
    b.close()
