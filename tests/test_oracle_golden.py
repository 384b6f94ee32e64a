"""The oracle (oracle/imex_oracle.c) and the host-side set-up against the
golden fixtures written by tests/golden/make_golden.py from the reference
solver modules compiled from their own sources.  All comparisons are
bit-for-bit."""
import numpy as np
import pytest

from conftest import golden_path
import cases
from oracle import oraclepy
from imex_sflow2d_dev_amd import init

FUNC_CASES = ["c1_48", "c2_64", "c3_64"]
RUN_CASES = ["c1_48", "c2_64", "c3_64"]


def _load(name):
    return dict(np.load(golden_path(name), allow_pickle=False))


def _eq(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, (what, a.shape, b.shape)
    if not np.array_equal(a, b, equal_nan=True):
        bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
        i = tuple(bad[0])
        raise AssertionError(f"{what}: {len(bad)} mismatches, first at {i}: {a[i]!r} vs {b[i]!r}")


@pytest.fixture(params=FUNC_CASES)
def func_case(request):
    cfg, _ = cases.CASES[request.param]()
    return cfg, _load(f"functions_{request.param}.npz")


def test_qc_to_qp(func_case):
    cfg, g = func_case
    o = oraclepy.Oracle(cfg)
    qp, pd = o.qc_to_qp(g["qc"])
    _eq(qp, g["qp"], "qp")
    _eq(pd, g["pdyn"], "p_dyn")


def test_qp_to_qc_and_mixt_var(func_case):
    cfg, g = func_case
    o = oraclepy.Oracle(cfg)
    _eq(o.qp_to_qc(g["qp"]), g["qp_to_qc"], "qp_to_qc")
    _eq(o.mixt_var(g["qp"]), g["mixt_var"], "mixt_var")


def test_fluxes_and_speeds(func_case):
    cfg, g = func_case
    o = oraclepy.Oracle(cfg)
    for d, lab in ((1, "x"), (2, "y")):
        _eq(o.fluxes(g["qc"], g["qp"], d), g["flux_" + lab], "flux_" + lab)
        _eq(o.speeds(g["qp"], d), g["speeds_" + lab], "speeds_" + lab)


def test_implicit_terms(func_case):
    cfg, g = func_case
    o = oraclepy.Oracle(cfg)
    _eq(o.implicit(g["qc"]), g["implicit"], "implicit")
    re, im = o.implicit_cplx(g["qc"], g["qim"])
    _eq(re, g["implicit_cplx_re"], "implicit complex re")
    _eq(im, g["implicit_cplx_im"], "implicit complex im")


def test_semi_implicit_and_explicit(func_case):
    cfg, g = func_case
    o = oraclepy.Oracle(cfg)
    _eq(o.semi_impl(g["qc"], g["qp"]), g["semi_impl"], "semi_impl")
    _eq(o.expl(g["bx"], g["by"], 3.0, g["qp"]), g["expl"], "expl")


def test_settling_and_mass_exchange(func_case):
    cfg, g = func_case
    o = oraclepy.Oracle(cfg)
    _eq(o.settling(g["s_diam"], g["s_rhos"], g["s_rhoc"], g["s_ikv"]), g["settling"], "settling")
    dep, eq = o.mass_exchange(g["qp"], float(g["mx_dt"]))
    _eq(dep, g["mx_dep"], "deposition")
    _eq(eq, g["mx_eq"], "mass exchange terms")


@pytest.fixture(params=RUN_CASES)
def run_case(request):
    cfg, B = cases.CASES[request.param]()
    g = _load(f"run_{request.param}.npz")
    _eq(B, g["B"], "bed")
    return cfg, B, g


def _static(g):
    st = {k[len("static_"):]: g[k] for k in g if k.startswith("static_")}
    st["h_source"] = float(g["h_source"])
    return st


def test_host_setup_matches_reference(run_case):
    """init.setup restates the reference's set-up: bed slopes, source ring,
    source thickness and the initial state, bit for bit."""
    cfg, B, g = run_case
    st, q0 = init.setup(cfg, B)
    for k, v in _static(g).items():
        if k == "h_source":
            assert st[k] == v
        elif k in st:
            _eq(st[k], v, k)
        else:
            assert not np.any(v), k
    _eq(q0, g["q0"], "q0")


def test_multistep_run(run_case):
    cfg, B, g = run_case
    o = oraclepy.Oracle(cfg, _static(g), g["q0"])
    n = len(g["dts"])
    dts, nc, err = o.run(n)
    assert err == 0
    _eq(dts, g["dts"], "dt sequence")
    _eq(nc, g["ncells"], "active cells")
    _eq(o.q(), g["q"], "q")
    f = o.fields()
    for k in ("hmax", "pdynmax", "mod_vel_max", "deposit"):
        _eq(f[k], g[k], k)
    assert f["t"] == float(g["t"])
    _eq(o.vuln().astype(np.int8), g["vuln"], "vulnerability")
    _eq(o.mask().astype(np.int8), g["mask"], "active set")
    dried, stale = o.counters()
    assert stale == 0
