"""Writes the transported-noise golden fixtures tests/golden/transport_<case>.npz
from the reference solver modules compiled from their own Fortran sources
(oracle/_ref/libimexref.so, built by `make -C oracle ref`), run with one
OpenMP thread.

    make -C oracle ref && python tests/golden/make_transport.py [case ...]

Cases (tests/cases.py TRANSPORT_CASES): the stochastic friction with the
noise transported by the flow (stoch_transport_flag; q1 * Z a conserved
variable after the solid classes) -- a one-class collapse (Voellmy), the
same with pore pressure and gas loss (whose mass exchange gives the noise a
source), and a two-class collapse.

Each fixture holds the inputs and the reference's outputs:
  q0, Z, the static fields, RUN_STEPS[case] driver steps: dts, active-cell
  counts, q, hmax, pdynmax, mod_vel_max;
  per-function vectors on states from the run (wet cells and 1 % random
  perturbations of them, f_qc): qc_to_qp, qp_to_qc, eval_fluxes (x, y),
  eval_mass_exchange_terms, solve_rk_step (from f_nt_qsi with q0 = f_qc);
and, beside them, the oracle's run of the same steps with correctly rounded
x**y, log, log1p and exp and states evaluated fresh (cr_*; cr_q and
cr_field_* only where they differ from the reference's q and field_*), the
arithmetic the HIP path computes.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from oracle import refpy  # noqa: E402
import cases  # noqa: E402

RUN_STEPS = {"c1t_64": 300, "cpot_64": 150, "c2t_48": 300}


def write(name):
    cfg, B, st, q0, Z = cases.TRANSPORT_CASES[name]()
    refpy.set_threads(1)
    run = refpy.RefRun(cfg, B)
    run.set_noise(Z)
    run.set_q(q0)
    static = run.static()
    steps = RUN_STEPS[name]
    dts, ncells = run.run(steps)
    out = dict(q0=q0, Z=Z, dts=dts, ncells=ncells, q=run.q(), steps=np.int32(steps))
    for k, v in static.items():
        out["static_" + k] = v
    for k, v in run.fields().items():
        out["field_" + k] = v
    q = run.q().reshape(-1, cfg.n_vars)
    wet = q[q[:, 0] > 0.0][::25]
    rng = np.random.default_rng(78)
    qc = np.concatenate([wet, wet * (1.0 + 1e-2 * rng.standard_normal(wet.shape))])
    qp, pd = run.qc_to_qp(qc)
    out.update(f_qc=qc, f_qp=qp, f_pdyn=pd, f_qc2=run.qp_to_qc(qp),
               f_flux_x=run.fluxes(qc, qp, 1), f_flux_y=run.fluxes(qc, qp, 2))
    dep, eq = run.mass_exchange(qp, 1.0e-3)
    out.update(f_me_dep=dep, f_me_eq=eq)
    dt = 1.0e-3
    qsi = qc * (1.0 + 1e-3 * rng.standard_normal(qc.shape))
    R = 1e-2 * qc * rng.standard_normal(qc.shape)
    out.update(f_nt_qsi=qsi, f_nt_R=R, f_nt_dt=np.float64(dt),
               f_nt_out=run.newton(dt, qsi, qc, R))
    assert np.array_equal(run.q(), out["q"])
    run.close()
    # the oracle with correctly rounded x**y, log, log1p and exp (the
    # rounding the HIP path computes) and fresh states, from the same q0:
    # what the GPU test compares bit for bit
    from oracle import oraclepy
    st2 = dict(static)
    st2["Z"] = Z
    o = oraclepy.Oracle(cfg, st2, q0)
    o.set_pow_mode(1)
    o.set_fresh(1)
    d, n, e = o.run(steps)
    # cr_q and the cr fields are stored only where they differ from the
    # reference's (read them with cr())
    out.update(cr_dts=d, cr_ncells=n, cr_err=np.int32(e),
               cr_stale_reads=np.int64(o.counters()[1]))
    if not np.array_equal(o.q(), out["q"]):
        out["cr_q"] = o.q()
    for k, v in o.fields().items():
        if not np.array_equal(v, out["field_" + k]):
            out["cr_field_" + k] = v
    np.savez_compressed(os.path.join(HERE, f"transport_{name}.npz"), **out)
    print(name, "written:", steps, "steps,", len(qc), "function states")


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(cases.TRANSPORT_CASES)):
        write(n)
