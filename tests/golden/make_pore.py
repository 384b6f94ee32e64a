"""Writes the pore-pressure golden fixtures tests/golden/pore_<case>.npz from
the reference solver modules compiled from their own Fortran sources
(oracle/_ref/libimexref.so, built by `make -C oracle ref`), run with one
OpenMP thread (with several, the reference's cell loops race on the module
variable hydraulic_permeability).

    make -C oracle ref && python tests/golden/make_pore.py [case ...]

Cases (tests/cases.py PORE_CASES): the physics of the four TEST_DAMBREAK
decks at 64 x 64 cells of 1 cm -- f_inhibit STATIC, DYNAMIC, OFF and STEP.

Each fixture holds the inputs and the reference's outputs:
  q0, the static fields, RUN_STEPS[case] driver steps: dts, active-cell
  counts, q, hmax, pdynmax, mod_vel_max, deposit;
  per-function vectors on states from the run (wet cells and 1 % random
  perturbations of them): qc_to_qp, eval_fluxes (x, y), eval_implicit_terms
  (real, and the complex step of each implicit column), eval_nh_semi_impl_
  terms, eval_mass_exchange_terms, solve_rk_step;
and, beside them, the oracle's run of the same steps with correctly rounded
x**y, log, log1p and exp and states evaluated fresh (cr_*), the arithmetic
the HIP path computes.
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

RUN_STEPS = {"cp_64": 150, "cpd_64": 150, "cpo_64": 120, "cps_64": 120}
EPS = 2.220446049250313e-16


def write(name):
    cfg, B, st, q0 = cases.PORE_CASES[name]()
    refpy.set_threads(1)
    run = refpy.RefRun(cfg, B)
    run.set_q(q0)
    static = run.static()
    steps = RUN_STEPS[name]
    dts, ncells = run.run(steps)
    out = dict(q0=q0, dts=dts, ncells=ncells, q=run.q(), steps=np.int32(steps))
    for k, v in static.items():
        out["static_" + k] = v
    for k, v in run.fields().items():
        out["field_" + k] = v
    # per-function vectors on the run's wet cells and perturbations of them
    q = run.q().reshape(-1, cfg.n_vars)
    wet = q[q[:, 0] > 0.0][::10]
    rng = np.random.default_rng(77)
    qc = np.concatenate([wet, wet * (1.0 + 1e-2 * rng.standard_normal(wet.shape))])
    qp, pd = run.qc_to_qp(qc)
    out.update(f_qc=qc, f_qp=qp, f_pdyn=pd,
               f_flux_x=run.fluxes(qc, qp, 1), f_flux_y=run.fluxes(qc, qp, 2),
               f_implicit=run.implicit(qc), f_semi=run.semi_impl(qc, qp))
    ip = 4 + cfg.n_solid
    for col in (1, 2, ip):
        qim = np.zeros_like(qc)
        qim[:, col] = cfg.n_vars * EPS * np.maximum(np.abs(qc[:, col]), 1e-3)
        re, im = run.implicit_cplx(qc, qim)
        out[f"f_cplx_im_{col}"] = qim
        out[f"f_cplx_re_{col}"] = re
        out[f"f_cplx_out_im_{col}"] = im
    dep, eq = run.mass_exchange(qp, 1.0e-3)
    out.update(f_me_dep=dep, f_me_eq=eq)
    dt = 1.0e-3
    qsi = qc * (1.0 + 1e-3 * rng.standard_normal(qc.shape))
    R = 1e-2 * qc * rng.standard_normal(qc.shape)
    out.update(f_nt_qsi=qsi, f_nt_q0=qc, f_nt_R=R, f_nt_dt=np.float64(dt),
               f_nt_out=run.newton(dt, qsi, qc, R))
    run.close()
    # the oracle with correctly rounded x**y, log, log1p and exp (the
    # rounding the HIP path computes) and fresh states, from the same q0:
    # what the GPU test compares bit for bit
    from oracle import oraclepy
    o = oraclepy.Oracle(cfg, static, q0)
    o.set_pow_mode(1)
    o.set_fresh(1)
    d, n, e = o.run(steps)
    out.update(cr_dts=d, cr_ncells=n, cr_err=np.int32(e), cr_q=o.q(),
               cr_stale_reads=np.int64(o.counters()[1]))
    for k, v in o.fields().items():
        out["cr_field_" + k] = v
    np.savez_compressed(os.path.join(HERE, f"pore_{name}.npz"), **out)
    print(name, "written:", steps, "steps,", len(qc), "function states")


if __name__ == "__main__":
    for n in (sys.argv[1:] or list(cases.PORE_CASES)):
        write(n)
