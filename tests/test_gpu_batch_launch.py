"""Launch shapes of a batch (imex_batch_create_form, imex_batch_set_launch):
members of either precision, the flux kernel on 256-thread blocks (two
passes per phase) or 512-thread blocks (one pass per phase, or two half
tiles per block when a member has more hot half tiles than blocks), and
steps replayed from captured HIP graphs or launched kernel by kernel.  Every
member equals the same run alone, bit for bit, after several run calls with
the host's own status reads and synchronisations between the graph replays
(the graphs are of 8 steps and of 1 step; 9-step calls use both)."""
import numpy as np
import pytest

from imex_sflow2d_dev_amd import init, workloads
from imex_sflow2d_dev_amd.ensemble import Batch, autotune, run_alone
from imex_sflow2d_dev_amd.solver import Solver

pytestmark = pytest.mark.gpu


def _bits(a, b, what):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{what}: {int((~same).sum())} mismatches"


def _members(n, t_end, precision, sides=(50, 90)):
    out = []
    for m in range(n):
        cfg, B = workloads.config4_member(m, min_side=sides[0], max_side=sides[1], t_end=t_end)
        cfg.precision = precision
        st, q0 = init.setup(cfg, B)
        out.append((cfg, st, q0))
    return out


def _radial_members(precision):
    out = []
    for n, dx in ((64, 20.0), (80, 16.0)):
        cfg, B = workloads.small_config2(nx=n, ny=n, dx=dx)
        cfg.precision = precision
        cfg.t_end = 0.4
        st, q0 = init.setup(cfg, B)
        out.append((cfg, st, q0))
    return out


def _check_batch(mem, n_solid, precision, flux_threads, graphs, blocks, chunk=9):
    print(f"batch: {len(mem)} members, n_solid={n_solid}, precision={precision}, "
          f"flux_threads={flux_threads}, graphs={graphs}, blocks={blocks}", flush=True)
    alone = []
    for cfg, st, q0 in mem:
        s = Solver(cfg, st, q0)
        run_alone(s, cfg.t_end)
        alone.append(s)
    b = Batch(n_solid=n_solid, blocks_per_member=blocks, precision=precision,
              flux_threads=flux_threads, graphs=graphs)
    inb = []
    for cfg, st, q0 in mem:
        s = Solver(cfg, st, q0)
        b.add(s, cfg.t_end)
        inb.append(s)
    calls = 0
    while True:
        done, _, _ = b.status()
        if done.all():
            break
        b.run(chunk)
        calls += 1
    assert calls >= 3
    for s in list(inb):
        b.remove(s)
    for i, (s, r) in enumerate(zip(inb, alone)):
        assert s.t == r.t, f"member {i}"
        _bits(s.q, r.q, f"member {i} q")
        fs, fr = s.fields(), r.fields()
        for k in ("hmax", "pdynmax", "mod_vel_max", "deposit"):
            _bits(fs[k], fr[k], f"member {i} {k}")
    b.close()
    for s in inb + alone:
        s.close()


@pytest.mark.parametrize("precision, flux_threads, graphs, blocks", [
    (0, 256, True, 8), (0, 512, True, 2), (0, 512, True, 64), (0, 512, False, 2),
    (1, 256, True, 8), (1, 512, True, 2), (1, 512, False, 64)])
def test_launch_shapes_match_runs_alone(precision, flux_threads, graphs, blocks):
    _check_batch(_members(4, 3.0, precision), 1, precision, flux_threads, graphs, blocks)


@pytest.mark.parametrize("precision, flux_threads", [(0, 512), (1, 256), (1, 512)])
def test_two_class_radial_members_launch_shapes(precision, flux_threads):
    _check_batch(_radial_members(precision), 2, precision, flux_threads, True, 4, chunk=17)


def test_fp32_member_rejected_by_fp64_batch():
    cfg, st, q0 = _members(1, 1.0, 1)[0]
    s = Solver(cfg, st, q0)
    b = Batch(n_solid=1)
    with pytest.raises(Exception):
        b.add(s, 1.0)
    b.close()
    s.close()


def test_autotune_keeps_members_exact():
    """The autotuner times launch shapes on the running batch; the members
    still equal their runs alone."""
    mem = _members(3, 2.0, 0)
    alone = []
    for cfg, st, q0 in mem:
        s = Solver(cfg, st, q0)
        run_alone(s, cfg.t_end)
        alone.append(s)
    b = Batch(n_solid=1)
    inb = [Solver(*m) for m in mem]
    for s, m in zip(inb, mem):
        b.add(s, m[0].t_end)
    best, table = autotune(b, blocks=(0, 8), flux_threads=(256, 512), steps=3)
    assert len(table) == 4 and best[1] in (256, 512)
    b.run_to_end(chunk=10)
    for s in list(inb):
        b.remove(s)
    for i, (s, r) in enumerate(zip(inb, alone)):
        assert s.t == r.t
        _bits(s.q, r.q, f"member {i} q")
    b.close()
