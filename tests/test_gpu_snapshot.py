"""Snapshots of output_solution's records (imex_snapshot_begin) against the
oracle, the in-place zeroing the reference applies to q and the deposit,
and the files of a run with solution outputs."""
import ctypes
import os

import numpy as np
import pytest

import cases
from imex_sflow2d_dev_amd import init, native, outputs
from imex_sflow2d_dev_amd.solver import Solver
from oracle import oraclepy, outputs_ref

pytestmark = pytest.mark.gpu

ALL = native.SNAP_Q | native.SNAP_Q2D | native.SNAP_DEP2D | native.SNAP_P2D | native.SNAP_FIELDS \
    | native.SNAP_ERO2D | native.SNAP_ERD2D


def _bufs(s, what):
    N = s.cfg.nx * s.cfg.ny
    return {k: native.PinnedArray(N * s.snapshot_cols(k)) for k in native.SNAP_KINDS if what & k}


@pytest.mark.parametrize("name, steps", [("c1_64", 200), ("c2_64", 200), ("c2e_64", 200),
                                         ("c1m_64", 200), ("c2m_64", 200)])
def test_snapshot_records_match_oracle(name, steps):
    print(f"snapshot {name} steps={steps}", flush=True)
    cfg, B = cases.CASES[name]()
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    bufs = None
    try:
        s.run_steps(steps)
        q = s.q
        # values the reference zeroes: tiny components of wet cells, a tiny
        # thickness, signed zeros
        wet = np.argwhere(q[:, :, 0] > 0)
        for n, (k, j) in enumerate(wet[:12]):
            q[k, j, 1 + n % 2] = 3e-21 * (-1) ** n
        k, j = wet[20]
        q[k, j, :] = 0.0
        q[k, j, 0] = 5e-21
        q[wet[30][0], wet[30][1], 2] = -0.0
        # a tiny thickness with momentum on dry cells beside the flow: the
        # .q_2d pass leaves them dry with other non-zero components, which the
        # next steps' active-set passes must visit
        dry = q[:, :, 0] <= 0
        nb = np.zeros_like(dry)
        nb[1:, :] |= ~dry[:-1, :]
        nb[:-1, :] |= ~dry[1:, :]
        nb[:, 1:] |= ~dry[:, :-1]
        nb[:, :-1] |= ~dry[:, 1:]
        edge = np.argwhere(dry & nb)
        assert len(edge) >= 8
        for n, (k, j) in enumerate(edge[::max(1, len(edge) // 8)][:8]):
            q[k, j, 0] = 5e-21
            q[k, j, 1 + n % 2] = 1e-3 * (-1) ** n
        t, d1, d2 = s.t, s.dt_old, s.dt_old_old
        s._upload(q)
        s.L.imex_set_time(s.h, ctypes.c_double(t), ctypes.c_double(d1), ctypes.c_double(d2))
        s.set_topography(B)
        f0 = s.fields()
        ero0, erd0 = s.erosion()
        N = cfg.nx * cfg.ny
        bufs = _bufs(s, ALL)
        s.snapshot_begin(ALL, {k: b.a for k, b in bufs.items()})
        s.snapshot_wait()
        o = oraclepy.Oracle(cfg)
        qz = outputs_ref.zero_tiny(q)
        X = cfg.x0 + (np.arange(cfg.nx) + 0.5) * cfg.dx
        Y = cfg.y0 + (np.arange(cfg.ny) + 0.5) * cfg.dx
        XX, YY = np.meshgrid(X, Y)
        xy = np.stack([XX.reshape(-1), YY.reshape(-1)], 1)
        q2d = bufs[native.SNAP_Q2D].a.reshape(N, -1)
        np.testing.assert_array_equal(q2d, np.concatenate([xy, qz.reshape(N, -1)], 1))
        assert not np.signbit(q2d[:, 2:][q2d[:, 2:] == 0]).any()
        # the state holds the zeroed values; the run's thin cell is now dry
        np.testing.assert_array_equal(s.q, qz)
        dep2d = bufs[native.SNAP_DEP2D].a.reshape(N, -1)
        np.testing.assert_array_equal(
            dep2d, np.concatenate([xy, f0["deposit"].reshape(cfg.n_solid, N).T], 1))
        want, depz, eroz = outputs_ref.p2d_records(cfg, qz, f0["deposit"], B, f0["hmax"],
                                                   f0["pdynmax"], f0["mod_vel_max"], o, ero0, erd0)
        got = bufs[native.SNAP_P2D].a.reshape(N, -1)
        np.testing.assert_array_equal(got, want)
        np.testing.assert_array_equal(bufs[native.SNAP_Q].a, qz.reshape(-1))
        ns = cfg.n_solid
        fld = bufs[native.SNAP_FIELDS].a.reshape(3 * ns + 4, N)
        np.testing.assert_array_equal(fld[3 * ns + 3], B.reshape(N))
        np.testing.assert_array_equal(fld[:ns], depz.reshape(ns, N))
        np.testing.assert_array_equal(fld[ns], f0["hmax"].reshape(N))
        np.testing.assert_array_equal(s.fields()["deposit"], depz)
        # erosion runs: the _erosion_ / _erodible_ rows before the .p_2d pass,
        # the zeroed erosion after it
        for kind, arr in ((native.SNAP_ERO2D, ero0), (native.SNAP_ERD2D, erd0)):
            np.testing.assert_array_equal(bufs[kind].a.reshape(N, -1),
                                          np.concatenate([xy, arr.reshape(ns, N).T], 1))
        np.testing.assert_array_equal(fld[ns + 3:2 * ns + 3], eroz.reshape(ns, N))
        np.testing.assert_array_equal(fld[2 * ns + 3:3 * ns + 3], erd0.reshape(ns, N))
        np.testing.assert_array_equal(s.erosion()[0], eroz)
        if cfg.erosion_coeff > 0.0:
            assert (erd0 < st["erodible"]).any() and (eroz > 0).any()
        # the run goes on from the zeroed state as the oracle does
        # a fresh start from the zeroed state, as after the upload
        o2 = oraclepy.Oracle(cfg, st, qz)
        o2.set_pow_mode(1)
        o2.set_time(t, d1, d2)
        if cfg.erosion_coeff > 0.0:
            o2.set_erodible(erd0)
        s.run_steps(30)
        o2.run(30)
        np.testing.assert_array_equal(s.q, o2.q())
    finally:
        if bufs:
            for b in bufs.values():
                b.free()
        s.close()


@pytest.mark.parametrize("name, topo", [("c2_64", False), ("c2e_64", True)])
def test_solution_files_of_a_run(tmp_path, name, topo):
    print(f"run with solution outputs {name} topo_change={topo}", flush=True)
    cfg, B = cases.CASES[name]()
    cfg.topo_change_flag = topo
    st, q0 = init.setup(cfg, B)
    s = Solver(cfg, st, q0)
    sol = outputs.SolutionOutput(s, "run", str(tmp_path), dt_output=0.4, B=B, n_threads=4)
    try:
        steps, _ = outputs.run_with_outputs(s, 1.2, solution=sol)
        q_end = s.q
        f_end = s.fields()
        ero_end, erd_end = s.erosion()
        b_end = s.topography() if topo else B
    finally:
        sol.close()
        s.close()
    names = sorted(os.listdir(tmp_path))
    for i in range(4):
        for suffix in (f"run_{i:04d}.q_2d", f"run_deposit_{i:04d}.q_2d", f"run_{i:04d}.p_2d"):
            assert suffix in names, (suffix, names)
    assert "restart.bin" in names and steps > 20
    if cfg.erosion_coeff > 0.0:
        for tag in ("erosion", "erodible") + (("B",) if topo else ()):
            assert f"run_{tag}_0003.q_2d" in names, (tag, names)
        if topo:
            b2d = np.loadtxt(tmp_path / "run_B_0003.q_2d")
            np.testing.assert_allclose(b2d[:, 2], b_end.reshape(-1), rtol=1e-11, atol=0)
    N, nv, ns = cfg.nx * cfg.ny, cfg.n_vars, cfg.n_solid
    # line layout: fixed-width records, a blank record after each row
    p2d = (tmp_path / "run_0003.p_2d").read_text().split("\n")
    assert len(p2d) == N + cfg.ny + 2 + 1 and p2d[cfg.nx] == "  " and p2d[-2] == "  "
    assert all(len(p2d[i]) == 15 * (19 + 4 * ns) for i in range(cfg.nx))
    q2d = np.loadtxt(tmp_path / "run_0003.q_2d")
    np.testing.assert_allclose(q2d[:, 2:], q_end.reshape(N, nv), rtol=1e-11, atol=0)
    # restart.bin: the reference's unformatted records, final state
    raw = (tmp_path / "restart.bin").read_bytes()
    off = 0
    recs = []
    while off < len(raw):
        n = int(np.frombuffer(raw, np.int32, 1, off)[0])
        recs.append(raw[off + 4:off + 4 + n])
        assert int(np.frombuffer(raw, np.int32, 1, off + 4 + n)[0]) == n
        off += 8 + n
    assert len(recs) == 10
    assert list(np.frombuffer(recs[0], np.int32)) == [cfg.nx, cfg.ny, nv, ns]
    np.testing.assert_array_equal(np.frombuffer(recs[2]), q_end.reshape(-1))
    np.testing.assert_array_equal(np.frombuffer(recs[5]), f_end["deposit"].reshape(-1))
    np.testing.assert_array_equal(np.frombuffer(recs[7]), f_end["hmax"].reshape(-1))
    np.testing.assert_array_equal(np.frombuffer(recs[3]), np.asarray(b_end).reshape(-1))
    np.testing.assert_array_equal(np.frombuffer(recs[6]), ero_end.reshape(-1))
    np.testing.assert_array_equal(np.frombuffer(recs[4]),
                                  np.transpose(erd_end, (1, 2, 0)).reshape(-1))
    t, dt = np.frombuffer(recs[1])
    assert t == pytest.approx(1.2, abs=1e-12)
