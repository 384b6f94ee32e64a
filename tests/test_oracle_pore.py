"""Pore pressure in the oracle (oracle/imex_oracle.c) against the reference:
the committed fixtures tests/golden/pore_<case>.npz written from the
reference solver modules built from their own sources
(tests/golden/make_pore.py), live runs of those modules where they are
built, the DGESV the reference links, and closed forms of the pore-pressure
terms.  The four cases carry the physics of the TEST_DAMBREAK decks
(f_inhibit STATIC, DYNAMIC, OFF, STEP) at 64 x 64 cells.  Bar: bit for bit
against the reference with the C library's functions (the oracle's default
mode); within 1e-10 between that mode and the correctly rounded one."""
import ctypes
import math

import numpy as np
import pytest

from conftest import golden_path
import cases
from oracle import oraclepy

EPS = 2.220446049250313e-16


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b, equal_nan=True):
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: {a[i]!r} vs {b[i]!r}")


def _fixture(name):
    return dict(np.load(golden_path(f"pore_{name}.npz"), allow_pickle=False))


def _static(g):
    return {k[len("static_"):]: v for k, v in g.items() if k.startswith("static_")}


def _dgesv_lib():
    """dgesv_ of the OpenBLAS build bundled with SciPy, the library the
    reference build links (oracle/Makefile)."""
    import glob
    import os
    import scipy
    d = os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs")
    libs = glob.glob(os.path.join(d, "libscipy_openblas-*.so"))
    if not libs:
        pytest.skip("SciPy's OpenBLAS not found")
    L = ctypes.CDLL(libs[0])
    return L.scipy_dgesv_


def test_dgesv_restatement_equals_the_linked_library():
    """oracle_dgesv (left-looking LU with fused dot products) against the
    library's DGESV on dense and on pore-structured 3 x 3 systems (the
    momentum rows do not depend on the pore column), bit for bit."""
    dgesv = _dgesv_lib()
    lib = oraclepy._lib()
    rng = np.random.default_rng(5)
    n = ctypes.c_int(3)
    one = ctypes.c_int(1)
    for structured in (False, True):
        for t in range(20000):
            a = (rng.random(9) - 0.5) * np.exp2(rng.integers(-10, 10, 9))
            b = (rng.random(3) - 0.5) * np.exp2(rng.integers(-10, 10, 3))
            if structured:
                a[6] = a[7] = 0.0      # column 3, rows 1 and 2 (column-major)
            a1, b1 = a.copy(), b.copy()
            ipiv = np.zeros(3, np.int32)
            info = ctypes.c_int(0)
            dgesv(ctypes.byref(n), ctypes.byref(one), a1.ctypes.data_as(ctypes.c_void_p),
                  ctypes.byref(n), ipiv.ctypes.data_as(ctypes.c_void_p),
                  b1.ctypes.data_as(ctypes.c_void_p), ctypes.byref(n), ctypes.byref(info))
            a2, b2 = a.copy(), b.copy()
            lib.oracle_dgesv(3, a2.ctypes.data_as(ctypes.c_void_p), b2.ctypes.data_as(ctypes.c_void_p))
            assert np.array_equal(b1, b2), (t, structured, b1, b2)


@pytest.mark.parametrize("name", sorted(cases.PORE_CASES))
def test_pore_functions_match_reference_fixture(name):
    cfg, B, st, q0 = cases.PORE_CASES[name]()
    g = _fixture(name)
    o = oraclepy.Oracle(cfg)
    qc, qp = g["f_qc"], g["f_qp"]
    qp2, pd2 = o.qc_to_qp(qc)
    _eq(qp2, qp, "qc_to_qp")
    _eq(pd2, g["f_pdyn"], "p_dyn")
    _eq(o.qp_to_qc(qp), o.qp_to_qc(qp2), "qp_to_qc")
    _eq(o.fluxes(qc, qp, 1), g["f_flux_x"], "flux x")
    _eq(o.fluxes(qc, qp, 2), g["f_flux_y"], "flux y")
    _eq(o.implicit(qc), g["f_implicit"], "implicit (real)")
    for col in (1, 2, 4 + cfg.n_solid):
        re, im = o.implicit_cplx(qc, g[f"f_cplx_im_{col}"])
        _eq(re, g[f"f_cplx_re_{col}"], f"implicit complex step, column {col}, real")
        _eq(im, g[f"f_cplx_out_im_{col}"], f"implicit complex step, column {col}, imag")
    _eq(o.semi_impl(qc, qp), g["f_semi"], "semi-implicit")
    dep, eq = o.mass_exchange(qp, 1.0e-3)
    _eq(dep, g["f_me_dep"], "deposition")
    _eq(eq, g["f_me_eq"], "mass exchange")
    dt = float(g["f_nt_dt"])
    for i in range(0, len(qc), 7):
        out, _ = o.newton(dt, g["f_nt_qsi"][i], g["f_nt_q0"][i], g["f_nt_R"][i])
        _eq(out, g["f_nt_out"][i], f"newton (3 x 3 DGESV), state {i}")


@pytest.mark.parametrize("name", sorted(cases.PORE_CASES))
def test_pore_run_matches_reference_fixture(name):
    cfg, B, st, q0 = cases.PORE_CASES[name]()
    g = _fixture(name)
    _eq(q0, g["q0"], "q0 (host set-up)")
    o = oraclepy.Oracle(cfg, _static(g), g["q0"])
    steps = int(g["steps"])
    dts, ncells, err = o.run(steps)
    _eq(dts, g["dts"], "dt sequence")
    _eq(ncells, g["ncells"], "active cells")
    _eq(o.q(), g["q"], "q")
    f = o.fields()
    for k in ("hmax", "pdynmax", "mod_vel_max"):
        _eq(f[k], g["field_" + k], k)
    assert err == 0
    # the pore pressure did something: p_exc fell below its start somewhere
    ip = 4 + cfg.n_solid
    wet = o.q()[..., 0] > 0
    p0 = g["q0"][..., ip] / np.where(g["q0"][..., 0] > 0, g["q0"][..., 0], 1.0)
    p1 = o.q()[..., ip] / np.where(wet, o.q()[..., 0], 1.0)
    assert np.max(p1[wet]) < np.max(p0)


@pytest.mark.slow
@pytest.mark.parametrize("name", sorted(cases.PORE_CASES))
def test_pore_multistep_against_live_reference(ref_lib, name):
    """300 driver steps of the reference (one OpenMP thread) and the oracle,
    bit for bit at every 100 steps."""
    cfg, B, st, q0 = cases.PORE_CASES[name]()
    ref_lib.set_threads(1)
    try:
        run = ref_lib.RefRun(cfg, B)
        run.set_q(q0)
        o = oraclepy.Oracle(cfg, run.static(), q0)
        for _ in range(3):
            d1, n1 = run.run(100)
            d2, n2, err = o.run(100)
            _eq(d2, d1, "dt")
            _eq(n2, n1, "active cells")
            _eq(o.q(), run.q(), "q")
            assert err == 0
        fr, fo = run.fields(), o.fields()
        for k in ("hmax", "pdynmax", "mod_vel_max"):
            _eq(fo[k], fr[k], k)
        run.close()
    finally:
        import os
        ref_lib.set_threads(os.cpu_count() or 1)


@pytest.mark.parametrize("name", ["cp_64", "cpd_64"])
def test_correctly_rounded_mode_within_tolerance(name):
    """The correctly rounded mode (the rounding the device computes: x**y,
    and cpow's log1p / log / exp, correctly rounded) against the C
    library's (the reference's) over 60 steps: 1e-10 relative L-infinity
    per variable over wet cells, the same wet mask."""
    cfg, B, st, q0 = cases.PORE_CASES[name]()
    g = _fixture(name)
    a = oraclepy.Oracle(cfg, _static(g), g["q0"])
    b = oraclepy.Oracle(cfg, _static(g), g["q0"])
    b.set_pow_mode(1)
    a.run(60)
    b.run(60)
    qa, qb = a.q(), b.q()
    wa, wb = qa[..., 0] > 0, qb[..., 0] > 0
    assert np.array_equal(wa, wb)
    scale = np.maximum(np.abs(qa[wa]).max(axis=0), 1e-300)
    rel = (np.abs(qa[wa] - qb[wa]) / scale).max()
    assert rel <= 1e-10, rel


def _smoothstep(N, x):
    """The generalised smoothstep S_N(x) = x^(N+1) sum_n C(N+n, n) C(2N+1, N-n) (-x)^n."""
    return x ** (N + 1) * sum(math.comb(N + n, n) * math.comb(2 * N + 1, N - n) * (-x) ** n
                              for n in range(N + 1))


def test_pore_source_closed_form():
    """The implicit pore-pressure source of a wet cell at rest equals
    -rho_m (pi/2)^2 D p_exc f_inhibit / h with D = k / (phi nu rho_g beta),
    k the Carman-Kozeny permeability phi^3 (d psi)^2 / (150 (1 - phi)^2),
    beta = 1 / (gamma p), gamma = c_p / (c_p - R), and f_inhibit
    1 - S_N((alpha - a_t) / (a_max - a_t)) (STATIC): a dropped term, a wrong
    power or a wrong coefficient of the smoothstep fails it."""
    cfg, B, st, q0 = cases.cp_64()
    o = oraclepy.Oracle(cfg)
    off_cfg, _, _, _ = cases.cp_64()
    off_cfg.f_inhibit_mode = "OFF"
    o_off = oraclepy.Oracle(off_cfg)
    ip = 4 + cfg.n_solid
    for alpha, h, pfrac in ((0.45, 0.2, 0.7), (0.5, 0.35, 0.9), (0.55, 0.05, 0.3), (0.3, 0.1, 0.5)):
        T = 300.0
        rho_g = cfg.pres / (cfg.sp_gas_const_a * T)
        rho_m = alpha * cfg.rho_s[0] + (1.0 - alpha) * rho_g
        qp = np.zeros(cfg.n_vars + 2)
        qp[0], qp[3], qp[4] = h, T, alpha * h
        pexc = pfrac * rho_m * cfg.grav * h
        qp[ip] = pexc
        qc = o.qp_to_qc(qp[None, :])[0]
        nh = o.implicit(qc[None, :])[0]
        nh_off = o_off.implicit(qc[None, :])[0]
        phi = 1.0 - alpha
        k = phi ** 3 * (cfg.diam_s[0] * 1.0) ** 2 / (150.0 * (1.0 - phi) ** 2)
        gamma = cfg.sp_heat_a / (cfg.sp_heat_a - cfg.sp_gas_const_a)
        D = k / (phi * cfg.kin_visc_a * rho_g / (gamma * cfg.pres))
        src_off = -rho_m * (math.pi / 2) ** 2 * D / h * pexc
        assert nh_off[ip] == pytest.approx(src_off, rel=1e-12)
        tt = min(1.0, max(0.0, (alpha - cfg.alpha_trans) / (cfg.maximum_solid_packing - cfg.alpha_trans)))
        f = 1.0 - _smoothstep(cfg.N_inh, tt)
        assert nh[ip] == pytest.approx(src_off * f, rel=1e-12, abs=1e-300)
        # no other row has an implicit term with mu(I) (model 11)
        assert np.count_nonzero(nh[:ip]) == 0


def test_gas_loss_closed_form():
    """The continuous-phase loss of the gas-loss branch,
    k / (nu rho_c) / max(h, 1e-5) (pi / 2) p_exc f_inhibit, enters the mass
    equation as -rho_c times it and the pore equation as that times p_exc."""
    cfg, B, st, q0 = cases.cp_64()
    off_cfg, _, _, _ = cases.cp_64()
    off_cfg.f_inhibit_mode = "OFF"
    o = oraclepy.Oracle(off_cfg)
    ip = 4 + cfg.n_solid
    alpha, h, T = 0.5, 0.2, 300.0
    rho_g = cfg.pres / (cfg.sp_gas_const_a * T)
    rho_m = alpha * cfg.rho_s[0] + (1.0 - alpha) * rho_g
    qp = np.zeros(cfg.n_vars + 2)
    qp[0], qp[3], qp[4] = h, T, alpha * h
    pexc = 0.4 * rho_m * cfg.grav * h
    qp[ip] = pexc
    dt = 1e-4
    dep, eq = o.mass_exchange(qp[None, :], dt)
    k = (1 - alpha) ** 3 * cfg.diam_s[0] ** 2 / (150.0 * alpha ** 2)
    loss = k / (cfg.kin_visc_a * rho_g) / h * 0.5 * math.pi * pexc
    loss = min(loss, h * max(0.0, cfg.maximum_solid_packing - alpha) / dt, h * (1 - alpha) / dt)
    assert eq[0, 0] == pytest.approx(-rho_g * loss, rel=1e-12)
    assert eq[0, ip] == pytest.approx(-rho_g * loss * pexc, rel=1e-12)
    assert eq[0, 1] == 0.0 and eq[0, 2] == 0.0


def test_pascal_coefficients_are_the_smoothstep():
    """precompute_pascal_coefficient's products C(-N-1, n) C(2N+1, N-n) are
    the coefficients of S_N: checked through f_inhibit at tt = 1/2 for N = 1..4
    (S_N(1/2) = 1/2 by symmetry) and at tt = 0.3 against the binomial form."""
    for N in (1, 2, 3, 4):
        cfg, B, st, q0 = cases.cp_64()
        cfg.N_inh = N
        o = oraclepy.Oracle(cfg)
        ip = 4 + cfg.n_solid
        off_cfg, _, _, _ = cases.cp_64()
        off_cfg.f_inhibit_mode = "OFF"
        o_off = oraclepy.Oracle(off_cfg)
        for tt in (0.5, 0.3):
            alpha = cfg.alpha_trans + tt * (cfg.maximum_solid_packing - cfg.alpha_trans)
            qp = np.zeros(cfg.n_vars + 2)
            qp[0], qp[3], qp[4], qp[ip] = 0.2, 300.0, alpha * 0.2, 2000.0
            qc = o.qp_to_qc(qp[None, :])[0]
            r = o.implicit(qc[None, :])[0, ip] / o_off.implicit(qc[None, :])[0, ip]
            assert r == pytest.approx(1.0 - _smoothstep(N, tt), rel=1e-9)
            if tt == 0.5:
                assert r == pytest.approx(0.5, rel=1e-9)


DAMBREAK_MODES = {"static": "STATIC", "dynamic": "DYNAMIC", "off": "OFF", "step": "STEP"}
# bound of the correctly rounded run against the reference's (the C
# library's pow, log, log1p, exp) at T_END: 1e-10, except for the DYNAMIC
# deck, where f_inhibit's transition fraction follows h**0.12 in every cell
# and step and the two roundings' runs part at step 116 and reach 2.1e-8
# relative (the GPU run is the correctly rounded one, bit for bit)
DAMBREAK_TOL = {"static": 1e-10, "dynamic": 5e-8, "off": 1e-10, "step": 1e-10}


@pytest.mark.parametrize("key", sorted(DAMBREAK_MODES))
def test_dambreak_decks_oracle_equals_reference(key):
    """The reference's TEST_DAMBREAK decks themselves (300 x 15 cells, their
    restart state, per-variable boundaries, to T_END = 0.3 s): the oracle
    with the C library's functions equals the reference's run bit for bit at
    every step (dt, active cells) and in q at the end; the correctly rounded
    run stays within 1e-10 of it (5e-8 for the DYNAMIC deck)."""
    cfg, B, st, q0 = cases.dambreak_deck(DAMBREAK_MODES[key])
    g = dict(np.load(golden_path(f"dambreak_{key}.npz"), allow_pickle=False))
    o = oraclepy.Oracle(cfg, st, q0)
    steps = int(g["steps"])
    dts, ncells, err = o.run(steps)
    _eq(dts, g["dts"], "dt sequence")
    _eq(ncells, g["ncells"], "active cells")
    _eq(o.q(), g["q"], "q")
    # (the running maxima are not compared: the harness takes the restart
    # state after its set-up, so its maxima leave out the initial state that
    # the driver's first epilogue, and the oracle's, include)
    assert err == 0
    t = cfg.t_start
    for d in dts:  # the driver's own accumulation t = t + dt reaches T_END
        t = t + float(d)
    assert t >= cfg.t_end
    q, qc = g["q"], g["cr_q"]
    w = q[..., 0] > 0.0
    assert np.array_equal(w, qc[..., 0] > 0.0)
    scale = np.maximum(np.abs(q[w]).max(axis=0), 1e-300)
    assert (np.abs(q[w] - qc[w]) / scale).max() <= DAMBREAK_TOL[key]
