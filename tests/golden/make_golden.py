"""Writes the golden fixtures of tests/golden/ from the reference solver
modules, compiled from their own Fortran sources into oracle/_ref by
oracle/Makefile (IEEE-preserving flags).

    make -C oracle ref && python tests/golden/make_golden.py [case ...]

Fixtures (numpy .npz, inputs and the reference's outputs):
  functions_<case>.npz  per-function vectors: qc_to_qp, qp_to_qc, mixt_var,
                        eval_fluxes and eval_local_speeds (x and y),
                        eval_implicit_terms (real and complex-step),
                        eval_nh_semi_impl_terms, eval_expl_terms,
                        settling_velocity, eval_mass_exchange_terms, on states
                        taken from a short run of the case plus hand-picked
                        branch states (dry, thin layer, solids only, T <= 0,
                        at rest);
  run_<case>.npz        a multi-step driver run: the static fields and q0 the
                        reference builds, and after N steps the dt sequence,
                        the active-cell counts, q, hmax, pdynmax, mod_vel_max,
                        deposit, the vulnerability flags and the active set.
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

RUN_STEPS = {"c1_48": 120, "c2_64": 150, "c3_64": 200}
FUNC_WARM = {"c1_48": 80, "c2_64": 100, "c3_64": 120}


def branch_states(cfg, eps_sing):
    """Hand-picked conservative states that reach each branch of qc_to_qp and
    the source terms."""
    nv, ns = cfg.n_vars, cfg.n_solid
    rows = []

    def st(q1, m1, m2, e, solids):
        r = np.zeros(nv)
        r[0], r[1], r[2], r[3] = q1, m1, m2, e
        r[4:4 + ns] = solids
        return r

    fr = np.array([0.03, 0.02, 0.01][:ns])
    rows.append(st(0.0, 0.0, 0.0, 0.0, 0.0 * fr))                      # dry
    rows.append(st(1e-17, 0.0, 0.0, 0.0, 0.0 * fr))                    # below EPSILON
    rows.append(st(0.5 * eps_sing, 1e-12, -2e-12, 1e-7, 0.5 * eps_sing * fr))   # thin layer
    rows.append(st(eps_sing, 3e-11, 1e-11, 3e-7, eps_sing * fr))
    rows.append(st(2.0 * eps_sing, 3e-11, 1e-11, 6e-7, 2.0 * eps_sing * fr))
    q1 = 50.0
    rows.append(st(q1, 0.0, 0.0, q1 * 1100.0 * 600.0, q1 * fr))        # at rest
    rows.append(st(q1, 120.0, -80.0, q1 * 1100.0 * 600.0, q1 * fr))
    rows.append(st(q1, 120.0, -80.0, -5.0, q1 * fr))                   # T <= 0 fallback
    solids = np.full(ns, q1 / ns)
    rows.append(st(q1, 30.0, 10.0, q1 * 1200.0 * 500.0, solids))       # solids only
    rows.append(st(1.0e4, -3.0e4, 2.0e4, 1.0e4 * 1100.0 * 900.0, 1.0e4 * fr))
    return np.array(rows)


def functions_fixture(name):
    cfg, B = cases.CASES[name]()
    run = refpy.RefRun(cfg, B)
    run.run(FUNC_WARM[name])
    q = run.q().reshape(-1, cfg.n_vars)
    rng = np.random.default_rng(7)
    wet = q[q[:, 0] > 0.0]
    pick = wet[rng.choice(len(wet), size=min(200, len(wet)), replace=False)]
    pert = pick * (1.0 + 1e-3 * rng.standard_normal(pick.shape))
    eps_sing = float(run.lib.imexref_eps_sing())
    qc = np.concatenate([pick, pert, branch_states(cfg, eps_sing)])
    n = len(qc)
    out = dict(qc=qc)
    qp, pd = run.qc_to_qp(qc)
    out.update(qp=qp, pdyn=pd)
    out["qp_to_qc"] = run.qp_to_qc(qp)
    out["mixt_var"] = run.mixt_var(qp)
    for d, lab in ((1, "x"), (2, "y")):
        out["flux_" + lab] = run.fluxes(qc, qp, d)
        out["speeds_" + lab] = run.speeds(qp, d)
    out["implicit"] = run.implicit(qc)
    qim = np.zeros_like(qc)
    hcs = cfg.n_vars * np.finfo(np.float64).eps
    qim[:, 1] = hcs * np.maximum(np.abs(qc[:, 1]), 1e-3)
    qim[n // 2:, 1] = 0.0
    qim[n // 2:, 2] = hcs * np.maximum(np.abs(qc[n // 2:, 2]), 1e-3)
    ore, oim = run.implicit_cplx(qc, qim)
    out.update(qim=qim, implicit_cplx_re=ore, implicit_cplx_im=oim)
    out["semi_impl"] = run.semi_impl(qc, qp)
    bx = 0.2 * rng.standard_normal(n)
    by = 0.2 * rng.standard_normal(n)
    out.update(bx=bx, by=by, expl=run.expl(bx, by, 3.0, qp))
    m = 400
    diam = 10.0 ** rng.uniform(-6.0, -1.5, m)
    rhos = rng.choice([800.0, 1500.0, 2500.0, 3000.0], m)
    rhoc = rng.uniform(0.3, 1.6, m)
    ikv = 1.0 / rng.uniform(1.0e-5, 3.0e-5, m)
    out.update(s_diam=diam, s_rhos=rhos, s_rhoc=rhoc, s_ikv=ikv,
               settling=run.settling(diam, rhos, rhoc, ikv))
    dep, eq = run.mass_exchange(qp, 0.05)
    out.update(mx_dt=np.array(0.05), mx_dep=dep, mx_eq=eq)
    run.close()
    np.savez_compressed(os.path.join(HERE, f"functions_{name}.npz"), **out)
    print(f"functions_{name}.npz: {n} states")


def run_fixture(name):
    cfg, B = cases.CASES[name]()
    run = refpy.RefRun(cfg, B)
    st = run.static()
    q0 = run.q()
    dts, nc = run.run(RUN_STEPS[name])
    f = run.fields()
    out = dict(B=B, q0=q0, dts=dts, ncells=nc, q=run.q(), hmax=f["hmax"],
               pdynmax=f["pdynmax"], mod_vel_max=f["mod_vel_max"], deposit=f["deposit"],
               t=np.array(f["t"]), vuln=run.vuln().astype(np.int8),
               mask=run.mask().astype(np.int8),
               h_source=np.array(st["h_source"]))
    for k in ("B_prime_x", "B_prime_y", "source_cell", "source_sides", "arc_perim",
              "arc_n_x", "arc_n_y"):
        out["static_" + k] = st[k]
    run.close()
    np.savez_compressed(os.path.join(HERE, f"run_{name}.npz"), **out)
    print(f"run_{name}.npz: {RUN_STEPS[name]} steps, t = {f['t']:.4f}, "
          f"active {nc[0]} -> {nc[-1]}")


def main():
    if not refpy.available("ieee"):
        sys.exit("oracle/_ref/libimexref.so is not built: make -C oracle ref")
    only = set(sys.argv[1:])            # case names; all when none are given
    for name in FUNC_WARM:
        if not only or name in only:
            functions_fixture(name)
    for name in RUN_STEPS:
        if not only or name in only:
            run_fixture(name)


if __name__ == "__main__":
    main()
