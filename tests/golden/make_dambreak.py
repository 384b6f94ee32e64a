"""Writes the fixtures of the reference's TEST_DAMBREAK decks
(tests/cases.py dambreak_deck): tests/golden/dambreak_q0.npz, the decks'
restart state (TESTS/TEST_DAMBREAK/TEST_STATIC_FINH/example_2D_0000.q_2d,
read as the reference's read_solution reads it, imex_sflow2d_dev_amd/
restart.py), and tests/golden/dambreak_<mode>.npz for each f_inhibit mode:
the run of the reference solver modules compiled from their own sources
(oracle/_ref/libimexref.so, one OpenMP thread) from that state to the
decks' T_END = 0.3 s -- dt sequence, active cells, q, running maxima -- and
beside it the oracle's run with correctly rounded x**y, log, log1p and exp
and fresh states (cr_*), the arithmetic the HIP path computes.

    make -C oracle ref && python tests/golden/make_dambreak.py [MODE ...]
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

DECK = "/root/reference/TESTS/TEST_DAMBREAK/TEST_STATIC_FINH/example_2D_0000.q_2d"
MODES = {"static": "STATIC", "dynamic": "DYNAMIC", "off": "OFF", "step": "STEP"}
MAX_STEPS = 20000


def write_q0():
    from imex_sflow2d_dev_amd import restart
    q0 = restart.read_q2d(DECK, 300, 15, 6, 1)
    np.savez_compressed(os.path.join(HERE, "dambreak_q0.npz"), q0=q0)


def write_mode(key):
    import cases
    from oracle import oraclepy, refpy
    cfg, B, st, q0 = cases.dambreak_deck(MODES[key])
    refpy.set_threads(1)
    run = refpy.RefRun(cfg, B)
    run.set_q(q0)
    static = run.static()
    # the driver loop runs while t < t_end (src/IMEX_SfloW2D.f90:395); the
    # harness steps on regardless, so it is called one step at a time and
    # t advanced as the driver advances it
    dts, ncells = [], []
    t = cfg.t_start
    while t < cfg.t_end and len(dts) < MAX_STEPS:
        d, n = run.run(1)
        dts.append(float(d[0]))
        ncells.append(int(n[0]))
        t = t + float(d[0])
    steps = len(dts)
    out = dict(steps=np.int32(steps), dts=np.array(dts), ncells=np.array(ncells, np.int32),
               q=run.q())
    for k, v in run.fields().items():
        out["field_" + k] = v
    run.close()
    o = oraclepy.Oracle(cfg, static, q0)
    o.set_pow_mode(1)
    o.set_fresh(1)
    d, n, e = o.run(steps)
    out.update(cr_dts=d, cr_ncells=n, cr_err=np.int32(e), cr_q=o.q(),
               cr_stale_reads=np.int64(o.counters()[1]))
    for k, v in o.fields().items():
        out["cr_field_" + k] = v
    np.savez_compressed(os.path.join(HERE, f"dambreak_{key}.npz"), **out)
    print(key, steps, "steps to t =", float(np.sum(dts)), "err", e, flush=True)


if __name__ == "__main__":
    if not os.path.exists(os.path.join(HERE, "dambreak_q0.npz")):
        write_q0()
    for k in (sys.argv[1:] or list(MODES)):
        write_mode(k)
