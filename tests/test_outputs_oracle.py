"""The oracle of the device-side outputs (oracle/outputs_ref.py) pinned on
the CPU: the driver's qp array against the reference solver's own, the
probe interpolation against the reference's compiled interp_2d_scalarB,
and the runout reductions against closed forms and brute force."""
import math

import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init, workloads
from oracle import oraclepy, outputs_ref


@pytest.mark.slow
@pytest.mark.parametrize("name, steps", [("c1_64", 150), ("c2_64", 200)])
def test_module_qp_is_qc_to_qp_of_the_state(ref_lib, name, steps):
    """The reference driver's qp array after its epilogues equals
    qc_to_qp(q) on wet cells and (0, T_ambient) elsewhere: the rule both the
    oracle and the device use for the outputs."""
    cfg, B = cases.CASES[name]()
    run = ref_lib.RefRun(cfg, B)
    try:
        run.run(steps)
        q, qp_ref = run.q(), run.qp()
    finally:
        run.close()
    o = oraclepy.Oracle(cfg)
    qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
    assert (q[:, :, 0] > 0).sum() > 50
    np.testing.assert_array_equal(qp, qp_ref)


@pytest.mark.slow
def test_interp_matches_reference_routine(ref_lib):
    cfg, B = cases.c1_48()
    run = ref_lib.RefRun(cfg, B)
    try:
        rng = np.random.default_rng(7)
        f = rng.standard_normal((cfg.ny, cfg.nx))
        x_comp = 0.5 * ((cfg.x0 + np.arange(cfg.nx) * cfg.dx) + (cfg.x0 + np.arange(1, cfg.nx + 1) * cfg.dx))
        y_comp = 0.5 * ((cfg.y0 + np.arange(cfg.ny) * cfg.dx) + (cfg.y0 + np.arange(1, cfg.ny + 1) * cfg.dx))
        L = cfg.nx * cfg.dx
        pts = [(rng.uniform(x_comp[0], L + 30.0), rng.uniform(-30.0, L + 30.0)) for _ in range(400)]
        # on cell centres, on the last centre, beyond the top and right edges
        pts += [(x_comp[j], y_comp[k]) for j, k in ((0, 0), (5, 7), (cfg.nx - 1, cfg.ny - 1),
                                                   (cfg.nx - 2, 3))]
        pts += [(L + 100.0, 100.0), (100.0, L + 100.0), (x_comp[3], -5.0)]
        xy = np.array(pts)
        want = run.interp(xy, f)
    finally:
        run.close()
    got = np.array([outputs_ref.interp_2d_scalarB(x_comp, y_comp, f, x, y) for x, y in xy])
    np.testing.assert_array_equal(got, want)
    assert np.count_nonzero(want) > 300


def _disc_state(cfg, xc, yc, R, h=2.0):
    """A uniform layer of thickness h on the cells whose centres lie within R
    of (xc, yc), at rest, T = 400 K, with solid fraction 0.05 per class."""
    o = oraclepy.Oracle(cfg)
    nx, ny, nv = cfg.nx, cfg.ny, cfg.n_vars
    x_comp = cfg.x0 + (np.arange(nx) + 0.5) * cfg.dx
    y_comp = cfg.y0 + (np.arange(ny) + 0.5) * cfg.dx
    X, Y = np.meshgrid(x_comp, y_comp)
    inside = (X - xc) ** 2 + (Y - yc) ** 2 <= R * R
    qp = np.zeros((int(inside.sum()), nv + 2))
    qp[:, 0] = h
    qp[:, 3] = 400.0
    for i in range(cfg.n_solid):
        qp[:, 4 + i] = 0.05 * h
    q = np.zeros((ny, nx, nv))
    q[inside] = o.qp_to_qc(qp)
    return q, inside, X, Y, o


def test_runout_of_a_disc_closed_form():
    cfg, _ = workloads.config1(nx=64, ny=64, dx=10.0)
    xc, yc = cfg.x0 + 32 * cfg.dx, cfg.y0 + 30 * cfg.dx   # a grid node: symmetric cell centres
    R = 150.0
    q, inside, X, Y, o = _disc_state(cfg, xc, yc, R)
    qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
    r = outputs_ref.runout_red(cfg, q, qp, np.zeros_like(inside), xc, yc, 0.0)
    # every cell of the disc is in hpos; the farthest centres from a node
    # lie at sqrt((a+0.5)^2 + (b+0.5)^2) dx for the largest such value <= R
    assert r["n_hpos"] == int(inside.sum()) == r["n_new"]
    d2 = [((a + 0.5) ** 2 + (b + 0.5) ** 2) * cfg.dx ** 2 for a in range(20) for b in range(20)
          if ((a + 0.5) ** 2 + (b + 0.5) ** 2) * cfg.dx ** 2 <= R * R]
    assert abs(r["dist_max"] - math.sqrt(max(d2))) < 1e-9
    # MAXLOC takes the first farthest cell in array order (k slowest)
    dist = np.where(inside, np.hypot(X - xc, Y - yc), 0.0)
    first = next(k * cfg.nx + j for k in range(cfg.ny) for j in range(cfg.nx)
                 if dist[k, j] == dist.max())
    assert r["imax"] == first
    # the barycentre of a symmetric disc is its centre
    assert abs(r["sum_xq1"] / r["sum_q1"] - xc) < 1e-9
    assert abs(r["sum_yq1"] / r["sum_q1"] - yc) < 1e-9
    # dist_x: farthest column from the centre; the first such cell in order
    assert abs(r["dist_x_max"] - (max(abs(X[inside] - xc)))) < 1e-9
    # nothing new against a mask holding the disc
    r2 = outputs_ref.runout_red(cfg, q, qp, inside, xc, yc, 0.0)
    assert r2["n_new"] == 0 and r2["n_hpos_old"] == r["n_hpos"]
    # a solid threshold above the fraction empties hpos: every dist is 0
    # and MAXLOC returns the first cell
    r3 = outputs_ref.runout_red(cfg, q, qp, inside, xc, yc, 1.0)
    assert r3["n_hpos"] == 0 and r3["imax"] == 0 and r3["dist_max"] == 0.0


def test_probe_values_of_a_uniform_layer():
    """Inside a uniform layer every interpolated field is the layer's own
    value; a point outside the grid reads the dry branch."""
    cfg, _ = workloads.config1(nx=48, ny=48, dx=10.0)
    cfg.settling_flag = False
    q, inside, X, Y, o = _disc_state(cfg, 240.0, 240.0, 200.0)
    qp = outputs_ref.module_qp(cfg, q, o.qc_to_qp)
    dep = np.zeros((cfg.n_solid, cfg.ny, cfg.nx))
    rows = outputs_ref.probes(cfg, q, qp, dep, [(241.3, 238.9), (10000.0, 10.0)])
    h, rho, T, u, v = rows[0, :5]
    assert abs(h - 2.0) < 1e-12 and abs(T - 400.0) < 1e-9 and u == 0.0 and v == 0.0
    assert abs(rho - q[24, 24, 0] / qp[24, 24, 0]) < 1e-9
    assert abs(rows[0, 5] - 0.05) < 1e-12 and rows[0, -1] == 0.0
    assert list(rows[1]) == [0.0, 0.0, cfg.T_ambient, 0.0, 0.0] + [0.0] * cfg.n_solid + [0.0]


def test_oracle_stops_at_t_end():
    """The C restatement's dt clip at a changed t_end (the output times of
    the device-driven schedule test) lands exactly on it."""
    cfg, B = cases.c1_48()
    st, q0 = init.setup(cfg, B)
    o = oraclepy.Oracle(cfg, st, q0)
    o.set_pow_mode(1)
    o.run(5)
    T = o.t + 0.05
    o.set_t_end(T)
    n = 0
    while o.t < T and n < 1000:
        o.run(1)
        n += 1
    assert o.t == T and n > 1


@pytest.mark.slow
@pytest.mark.parametrize("name, steps", [("c1m_64", 200), ("c2m_64", 200)])
def test_p2d_mu_i_columns_against_reference_friction(ref_lib, name, steps):
    """The .p_2d mu_eff column of the mu(I) models (src/inpout_2d.f90:
    6277-6335) against the friction the reference's own semi-implicit terms
    apply (src/constitutive_2d.f90:4384-4537): with no slope correction both
    evaluate mu(I) of the same inertial number, so the momentum source is
    -mu_eff * sigma_eff * u / |u|, sigma_eff = MAX(rho_m g' h, EPSILON).  At
    rest mu_eff is mu_s and the inertial number 0."""
    cfg, B = cases.CASES[name]()
    run = ref_lib.RefRun(cfg, B)
    try:
        run.run(steps)
        q = run.q()
        N = cfg.nx * cfg.ny
        qc = q.reshape(N, cfg.n_vars)
        qp, _ = run.qc_to_qp(qc)
        si = run.semi_impl(qc, qp)
        mx = run.mixt_var(qp)
    finally:
        run.close()
    o = oraclepy.Oracle(cfg)
    z = np.zeros((cfg.n_solid, cfg.ny, cfg.nx))
    rec, _, _ = outputs_ref.p2d_records(cfg, q, z, np.zeros((cfg.ny, cfg.nx)), z[0], z[0],
                                        z[0], o)
    mu_eff, inertial = rec[:, -5], rec[:, -1]
    nv = cfg.n_vars
    h, u, v = qp[:, 0], qp[:, nv], qp[:, nv + 1]
    speed = np.sqrt(u * u + v * v)
    eps = np.finfo(np.float64).eps
    ens = np.maximum(mx[:, 1] * mx[:, 3] * np.maximum(h, eps), eps)
    sel = (h > 1e-3) & (np.abs(u) > 0.1 * speed) & (speed > 1e-8)
    assert sel.sum() > 100
    mu_ref = -si[sel, 1] / (u[sel] / speed[sel]) / ens[sel]
    np.testing.assert_allclose(mu_eff[sel], mu_ref, rtol=1e-12, atol=0)
    # mu(I) rises from mu_s with the inertial number; the runs reach I well
    # above I_0 / 1000
    assert (mu_eff[sel] > cfg.mu_s).all() and inertial[sel].max() > 1e-4 * cfg.I_0
    rest = (h > 0) & (speed < eps)
    np.testing.assert_array_equal(mu_eff[rest], cfg.mu_s)
    np.testing.assert_array_equal(inertial[rest], 0.0)
    assert (mu_eff[h <= 0] == 0).all() and (inertial[h <= 0] == 0).all()
