"""Seeded synthetic workloads (numpy PCG64, seed = 1000*config + member).

Every bed elevation is returned at the cell centres; the reference reads an
ESRI grid and interpolates it bilinearly to the centres
(src/geometry_2d.f90:336-362), which returns the pixel values exactly when the
pixels sit on the centres, and clamps it to >= 0 (src/geometry_2d.f90:350).
"""
from __future__ import annotations

import math

import numpy as np

from .config import RunConfig


def _rng(config_id: int, member: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64(1000 * config_id + member))


def cell_centres(cfg: RunConfig):
    """x_comp, y_comp as the reference builds them (src/geometry_2d.f90:312-330)."""
    j = np.arange(cfg.nx + 1, dtype=np.float64)
    k = np.arange(cfg.ny + 1, dtype=np.float64)
    x_stag = cfg.x0 + j * cfg.dx
    y_stag = cfg.y0 + k * cfg.dx
    x_comp = 0.5 * (x_stag[:-1] + x_stag[1:])
    y_comp = 0.5 * (y_stag[:-1] + y_stag[1:])
    return x_comp, y_comp, x_stag, y_stag


def fractal_surface(n: int, dx: float, beta: float, rms: float,
                    rng: np.random.Generator) -> np.ndarray:
    """Spectral synthesis: random phases with |A(k)|^2 ~ |k|^beta, rescaled to
    the requested RMS.  Returned as [k, j] (y rows, x columns)."""
    kx = np.fft.fftfreq(n, d=dx)
    ky = np.fft.fftfreq(n, d=dx)
    kk = np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)
    kk[0, 0] = 1.0
    amp = kk ** (beta / 2.0)
    amp[0, 0] = 0.0
    phase = rng.uniform(0.0, 2.0 * np.pi, size=(n, n))
    spec = amp * np.exp(1j * phase)
    surf = np.real(np.fft.ifft2(spec))
    surf -= surf.mean()
    surf *= rms / np.sqrt(np.mean(surf ** 2))
    return surf


def config1(nx: int = 256, ny: int = 256, dx: float = 10.0, seed_member: int = 0):
    """Plain method, small, FP64: 5-degree planar slope plus 8 Gaussian bumps,
    cylinder of radius 300 m and thickness 20 m at rest upslope, one solid
    class, Voellmy mu=0.1 xi=1000, KT + minmod, IMEX RK2, CFL 0.24,
    transmissive boundaries.  Returns (cfg, B[k, j])."""
    cfg = RunConfig(nx=nx, ny=ny, dx=dx, name="C1")
    rng = _rng(1, seed_member)
    xc, yc, _, _ = cell_centres(cfg)
    lx, ly = nx * dx, ny * dx
    X, Y = np.meshgrid(xc, yc)               # [k, j]
    B = 10.0 + (lx - X) * math.tan(math.radians(5.0))
    for _ in range(8):
        a = rng.uniform(5.0, 30.0)
        s = rng.uniform(50.0, 200.0)
        cx = rng.uniform(0.0, lx)
        cy = rng.uniform(0.0, ly)
        B = B + a * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * s * s))
    cfg.x_collapse = 0.25 * lx
    cfg.y_collapse = 0.5 * ly
    cfg.r_collapse = min(300.0, 0.2 * min(lx, ly))
    cfg.validate()
    return cfg, np.ascontiguousarray(B)


def config2(nx: int = 2048, ny: int = 2048, dx: float = 5.0, seed_member: int = 0):
    """Plain drop-in at production size, FP64: spectral fractal DEM (power-law
    exponent -2.5, RMS 20 m) on a 15-degree cone; lateral radial source of
    radius 100 m at the summit, mass flux 1e7 kg/s for 120 s, inflow 800 K,
    two solid classes (1e-3 and 1e-4 m, 2500 kg/m3, fractions 0.03/0.02),
    sedimentation and air entrainment on.  Returns (cfg, B[k, j])."""
    cfg = RunConfig(nx=nx, ny=ny, dx=dx, name="C2")
    rng = _rng(2, seed_member)
    xc, yc, _, _ = cell_centres(cfg)
    lx, ly = nx * dx, ny * dx
    X, Y = np.meshgrid(xc, yc)
    # The summit (cone apex and source centre) is kept off the grid's symmetry
    # points: a flow that is exactly mirror-symmetric about a cell face gives a
    # mass flux of exactly zero there, and the reference's x-direction solid
    # flux cap then divides by (-Inf)+(+Inf) (src/solver_2d.f90:2903-2912).
    cx, cy = 0.5 * lx + 0.31 * dx, 0.5 * ly - 0.43 * dx
    r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
    n = max(nx, ny)
    frac = fractal_surface(n, dx, -2.5, 20.0, rng)[:ny, :nx]
    B = 200.0 + np.max(r) * math.tan(math.radians(15.0)) \
        - r * math.tan(math.radians(15.0)) + frac
    B = np.maximum(B, 0.0)
    cfg.n_solid = 2
    cfg.rho_s = [2500.0, 2500.0]
    cfg.diam_s = [1.0e-3, 1.0e-4]
    cfg.sp_heat_s = [1200.0, 1200.0]
    cfg.settling_flag = True
    cfg.entrainment_flag = True
    cfg.collapsing_volume_flag = False
    cfg.alphas_collapse = [0.0, 0.0]
    cfg.radial_source_flag = True
    cfg.x_source = cx
    cfg.y_source = cy
    cfg.r_source = 100.0
    cfg.vel_source = 20.0
    cfg.T_source = 800.0
    cfg.mfr_source = 1.0e7
    # with MFR_SOURCE given the duration is forced to 1e14/MFR
    # (src/inpout_2d.f90:4435-4437); TIME_PARAM(4) is the cut-off time
    cfg.time_param = [1.0e7, 1.0e14 / cfg.mfr_source, 0.0, 120.0]
    cfg.alphas_source = [0.03, 0.02]
    cfg.validate()
    return cfg, np.ascontiguousarray(B)


def config4_member(member: int, min_side: int = 128, max_side: int = 1024,
                   t_end: float = 30.0, n_solid: int = 1):
    """One member of the batched ensemble: a grid side log-uniform in
    [min_side, max_side] (dx = 10 m), a seeded fractal DEM (exponent -2.5,
    RMS 10 m) on a 5-degree slope, and per member an initial volume
    U[1e5, 1e7] m3 realised as a cylinder of thickness 20 m (radius
    sqrt(V / (pi 20))) at rest upslope, mu U[0.05, 0.3], xi U[200, 2000],
    initial T U[400, 900] K and particle fraction U[0.01, 0.1]; the config-1
    physics otherwise (one class, no settling).  Each member runs to t_end.
    Returns (cfg, B[k, j])."""
    rng = _rng(4, member)
    n = int(round(math.exp(rng.uniform(math.log(min_side), math.log(max_side)))))
    dx = 10.0
    cfg = RunConfig(nx=n, ny=n, dx=dx, name=f"C4m{member}")
    xc, yc, _, _ = cell_centres(cfg)
    lx = n * dx
    X, _ = np.meshgrid(xc, yc)
    frac = fractal_surface(n, dx, -2.5, 10.0, rng)
    B = 50.0 + (lx - X) * math.tan(math.radians(5.0)) + frac
    B = np.maximum(B, 0.0)
    vol = rng.uniform(1.0e5, 1.0e7)
    cfg.h_collapse = 20.0
    cfg.r_collapse = min(math.sqrt(vol / (math.pi * cfg.h_collapse)), 0.2 * lx)
    # off the cell faces' symmetry lines (see config2)
    cfg.x_collapse = 0.3 * lx + 0.31 * dx
    cfg.y_collapse = 0.5 * lx - 0.43 * dx
    cfg.mu = rng.uniform(0.05, 0.3)
    cfg.xi = rng.uniform(200.0, 2000.0)
    cfg.T_collapse = rng.uniform(400.0, 900.0)
    cfg.alphas_collapse = [rng.uniform(0.01, 0.1)]
    if n_solid > 1:
        # the particle fraction split over n_solid classes of the radial-source
        # workload's sizes (1e-3, 1e-4 m; 1e-2 m first for three), with settling
        diam = {2: [1.0e-3, 1.0e-4], 3: [1.0e-2, 1.0e-3, 1.0e-4]}[n_solid]
        cfg.n_solid = n_solid
        cfg.rho_s = [2500.0] * n_solid
        cfg.diam_s = diam
        cfg.sp_heat_s = [1200.0] * n_solid
        cfg.alphas_collapse = [cfg.alphas_collapse[0] / n_solid] * n_solid
        cfg.alphas_source = [0.0] * n_solid
        cfg.settling_flag = True
    cfg.t_end = t_end
    cfg.validate()
    return cfg, np.ascontiguousarray(B)


def config4(n_members: int = 256, **kw):
    """The batched ensemble: [(cfg, B), ...] for members 0..n_members-1."""
    return [config4_member(m, **kw) for m in range(n_members)]


def small_config2(nx: int = 96, ny: int = 96, dx: float = 20.0):
    """A CPU-sized version of the production workload (same physics and
    source, coarser and smaller grid) for parity tests against the oracle."""
    cfg, B = config2(nx=nx, ny=ny, dx=dx)
    cfg.name = "C2small"
    cfg.r_source = 3.2 * dx
    return cfg, B


def config3(nx: int = 4096, ny: int = 4096, dx: float = 2.5, seed_member: int = 0):
    """The config-1 physics (no sedimentation) on a larger grid with three
    solid classes (1e-2/1e-3/1e-4 m).  The per-class fractions are not fixed
    by the workload description: 0.02/0.02/0.01 is used, and the collapsing
    volume is scaled with the domain (radius 3 % of the domain side)."""
    cfg, B = config1(nx=nx, ny=ny, dx=dx, seed_member=seed_member)
    cfg.name = "C3"
    cfg.n_solid = 3
    cfg.rho_s = [2500.0, 2500.0, 2500.0]
    cfg.diam_s = [1.0e-2, 1.0e-3, 1.0e-4]
    cfg.sp_heat_s = [1200.0, 1200.0, 1200.0]
    cfg.alphas_collapse = [0.02, 0.02, 0.01]
    cfg.alphas_source = [0.0, 0.0, 0.0]
    cfg.r_collapse = 0.03 * nx * dx
    cfg.validate()
    return cfg, B


WORKLOADS = {
    "c1": config1,
    "c2": config2,
    "c2-small": small_config2,
    "c3": config3,
}


def by_name(name: str):
    """(RunConfig, bed) of a named synthetic workload."""
    try:
        return WORKLOADS[name]()
    except KeyError:
        raise ValueError(f"unknown workload {name!r}; known: {sorted(WORKLOADS)}") from None
