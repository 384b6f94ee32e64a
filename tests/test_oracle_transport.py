"""The transported noise (stoch_transport_flag) in the oracle
(oracle/imex_oracle.c) against the reference: the committed fixtures
tests/golden/transport_<case>.npz written from the reference solver modules
built from their own sources (tests/golden/make_transport.py), live runs of
those modules where they are built, and what the mathematics fixes -- a
noise that is the same constant everywhere stays the same multiple of the
mass, through the advection and the mass exchange alike.  Bar: bit for bit
against the reference."""
import numpy as np
import pytest

from conftest import golden_path
import cases
from oracle import oraclepy


def _eq(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b, equal_nan=True):
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: {a[i]!r} vs {b[i]!r}")


def _fixture(name):
    return dict(np.load(golden_path(f"transport_{name}.npz"), allow_pickle=False))


def _static(g):
    st = {k[len("static_"):]: v for k, v in g.items() if k.startswith("static_")}
    st["Z"] = g["Z"]
    return st


@pytest.mark.parametrize("name", sorted(cases.TRANSPORT_CASES))
def test_transport_functions_match_reference_fixture(name):
    cfg, B, st, q0, Z = cases.TRANSPORT_CASES[name]()
    g = _fixture(name)
    o = oraclepy.Oracle(cfg)
    qc, qp = g["f_qc"], g["f_qp"]
    qp2, pd2 = o.qc_to_qp(qc)
    _eq(qp2, qp, "qc_to_qp")
    _eq(pd2, g["f_pdyn"], "p_dyn")
    _eq(o.qp_to_qc(qp), g["f_qc2"], "qp_to_qc")
    _eq(o.fluxes(qc, qp, 1), g["f_flux_x"], "flux x")
    _eq(o.fluxes(qc, qp, 2), g["f_flux_y"], "flux y")
    dep, eq = o.mass_exchange(qp, 1.0e-3)
    _eq(dep, g["f_me_dep"], "deposition")
    _eq(eq, g["f_me_eq"], "mass exchange")
    dt = float(g["f_nt_dt"])
    for i in range(0, len(qc), 3):
        out, _ = o.newton(dt, g["f_nt_qsi"][i], qc[i], g["f_nt_R"][i])
        _eq(out, g["f_nt_out"][i], f"newton, state {i}")
    # the noise variable is the mass times Z, and its fluxes the mass
    # fluxes times Z where the face state is wet
    ist = 4 + cfg.n_solid
    wet = qc[:, 0] > 0.0
    np.testing.assert_allclose(qp[wet, ist] * qc[wet, 0], qc[wet, ist], rtol=1e-14, atol=0)
    if cfg.pore_pressure_flag:
        # the gas loss gives the noise a source: eqns_term(1) * Z
        assert np.any(eq[:, ist] != 0.0)
        np.testing.assert_array_equal(eq[:, ist], eq[:, 0] * qp[:, ist])


@pytest.mark.parametrize("name", sorted(cases.TRANSPORT_CASES))
def test_transport_run_matches_reference_fixture(name):
    cfg, B, st, q0, Z = cases.TRANSPORT_CASES[name]()
    g = _fixture(name)
    _eq(q0, g["q0"], "q0 (host set-up)")
    _eq(Z, g["Z"], "Z (host burn-in)")
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
    # the correctly rounded run (the HIP path's rounding) is the same run
    _eq(g["cr_dts"], g["dts"], "correctly rounded dt sequence")
    assert "cr_q" not in g and int(g["cr_err"]) == 0
    # the noise moved with the flow
    ist = 4 + cfg.n_solid
    assert not np.array_equal(o.q()[..., ist], g["q0"][..., ist])


@pytest.mark.parametrize("name", ["c1t_64", "cpot_64"])
def test_constant_noise_stays_proportional_to_the_mass(name):
    """With Z the same constant c in every cell, q1 * Z starts as c q1; its
    flux u (c q1) and its mass-exchange source eqns_term(1) c are c times
    those of q1, so after any number of steps q_Z = c q1 up to the
    roundings of the two products (relative 1e-12 of the largest c q1 over
    100 steps), and dry cells hold 0 in both."""
    cfg, B, st, q0, Z = cases.TRANSPORT_CASES[name]()
    c = 37.25
    Zc = np.full_like(Z, c)
    q0c = q0.copy()
    ist = 4 + cfg.n_solid
    q0c[..., ist] = q0c[..., 0] * c
    stt = dict(st)
    stt["Z"] = Zc
    o = oraclepy.Oracle(cfg, stt, q0c)
    dts, n, err = o.run(100)
    assert err == 0
    q = o.q()
    scale = np.max(np.abs(c * q[..., 0]))
    assert np.max(np.abs(q[..., ist] - c * q[..., 0])) <= 1e-12 * scale
    assert np.all(q[q[..., 0] == 0.0][:, ist] == 0.0)
    # and the mass itself moved (the check is not on a frozen state)
    assert not np.array_equal(q[..., 0], q0c[..., 0])


@pytest.mark.slow
@pytest.mark.parametrize("name", sorted(cases.TRANSPORT_CASES))
def test_transport_multistep_against_live_reference(ref_lib, name):
    """300 driver steps of the reference (one OpenMP thread) and the oracle
    from the same q0 and Z, bit for bit at every 100 steps."""
    cfg, B, st, q0, Z = cases.TRANSPORT_CASES[name]()
    ref_lib.set_threads(1)
    run = ref_lib.RefRun(cfg, B)
    run.set_noise(Z)
    run.set_q(q0)
    stt = run.static()
    stt["Z"] = Z
    o = oraclepy.Oracle(cfg, stt, q0)
    for _ in range(3):
        d1, n1 = run.run(100)
        d2, n2, err = o.run(100)
        _eq(d2, d1, "dt")
        _eq(n2, n1, "active cells")
        _eq(o.q(), run.q(), "q")
        assert err == 0
    run.close()
