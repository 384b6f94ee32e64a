"""The FP32 form against the FP64 drop-in on the same run: the quantities
the precision form is compared on -- runout area, maximum thickness, deposit
and conserved mass -- from the final state and output fields of each."""
from __future__ import annotations

import copy
import time

import numpy as np

from .solver import Solver


def run_stats(s: Solver) -> dict:
    """Runout area (cells the flow reached, hmax > 1e-3 m), maximum
    thickness, deposit (volume and maximum thickness, over all classes) and
    mass (in the flow plus deposited)."""
    cfg = s.cfg
    f = s.fields()
    q = s.q
    area = cfg.dx * cfg.dx
    dep = f["deposit"]                       # [class, k, j] thickness
    dep_tot = dep.sum(axis=0)
    rho_s = np.asarray(cfg.rho_s[:cfg.n_solid])
    dep_mass = float((dep.sum(axis=(1, 2)) * rho_s).sum()) * area
    return {
        "runout_area_m2": float(area * np.count_nonzero(f["hmax"] > 1.0e-3)),
        "max_thickness_m": float(f["hmax"].max()),
        "deposit_volume_m3": float(area * dep_tot.sum()),
        "max_deposit_m": float(dep_tot.max()),
        "flow_mass_kg": float(area * q[:, :, 0].sum()),
        "total_mass_kg": float(area * q[:, :, 0].sum() + dep_mass),
    }


def compare(cfg, static, q0, steps: int, chunk: int = 500, log=None) -> dict:
    """The same run in FP64 and in the FP32 form, each timed over `steps`
    device-driven steps; returns both stats, the relative differences and
    the two rates."""
    say = log or (lambda *a: None)
    out = {}
    for prec in (0, 1):
        c = copy.deepcopy(cfg)
        c.precision = prec
        s = Solver(c, static, q0, raise_on_error=False)
        s.run_steps(0)
        cells = 0
        t0 = time.perf_counter()
        done = 0
        while done < steps:
            n = min(chunk, steps - done)
            s.run_steps(n, record_dt=False)
            done += n
        wall = time.perf_counter() - t0
        a, _ = s.counters()
        key = "fp64" if prec == 0 else "fp32"
        out[key] = dict(run_stats(s), wall_s=wall, t_end_s=s.t, err_bits=s.err,
                        cell_stage_updates_per_s=2.0 * a / wall)
        say(f"{key}: {wall:.3f} s, t = {s.t:.4f}")
        s.close()
    rel = {}
    for k in ("runout_area_m2", "max_thickness_m", "deposit_volume_m3", "max_deposit_m",
              "flow_mass_kg", "total_mass_kg"):
        ref = out["fp64"][k]
        rel[k] = abs(out["fp32"][k] - ref) / abs(ref) if ref != 0.0 else abs(out["fp32"][k])
    out["relative_difference"] = rel
    out["speedup_fp32"] = out["fp64"]["wall_s"] / out["fp32"]["wall_s"]
    return out
