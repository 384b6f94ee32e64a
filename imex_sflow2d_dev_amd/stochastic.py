"""Host-side set-up of the stochastic friction: the cells' values Z of the
Ornstein-Uhlenbeck process after the burn-in of getSteadyStateZ
(src/stochastic_module.f90:79-132), which the reference runs once, before
the first step, when it restarts from an ASCII solution.

The process: Z = 0, then ceil(10 tau / (0.1 tau)) Euler-Maruyama steps of
dt = 0.1 tau (EulerMaruyamaScheme, :367-375)
    Z <- Z - (dt / tau) Z + sigma sqrt(2 dt / tau) N,
N standard normal noise over the grid (GaussianNoise, Box-Muller, :240-257),
smoothed by a normalised Gaussian kernel of width length_spatial_corr when
that exceeds the cell size (genConvolutionKernel / convolve_2d, :135-182,
:377-430), sigma = std_max for the friction laws on this path
(getSigmaNoise, :340-358), and the sym_noise map |Z|**noise_pow_val (or its
negative) after every step.  The random numbers are numpy's, not the
Fortran runtime's: the values are a sample of the same process, not the
reference's sample.
"""
from __future__ import annotations

import math

import numpy as np


def gaussian_kernel(length_spatial_corr: float, cell_size: float) -> np.ndarray:
    """genConvolutionKernel (src/stochastic_module.f90:135-182): a square
    kernel of ceil(L / dx) cells sampling exp(-r^2 / (2 L^2)) at the cell
    centres about the kernel's centre, normalised to sum 1."""
    n = int(math.ceil(length_spatial_corr / cell_size))
    c = cell_size * n / 2.0
    x = cell_size * np.arange(n) + 0.5 * cell_size
    dx = (x - c) / length_spatial_corr
    k = np.exp(-0.5 * (dx[None, :] ** 2 + dx[:, None] ** 2))
    return k / k.sum()


def _convolve(noise: np.ndarray, kern: np.ndarray) -> np.ndarray:
    """convolve_2d (src/stochastic_module.f90:377-430): zero padding of
    (n - 1) // 2 cells before and the rest after, kernel centred on the cell."""
    n = kern.shape[0]
    lo = (n - 1) // 2
    hi = n - 1 - lo
    pad = np.pad(noise, ((lo, hi), (lo, hi)))
    out = np.zeros_like(noise)
    for a in range(n):
        for b in range(n):
            out += kern[a, b] * pad[a:a + noise.shape[0], b:b + noise.shape[1]]
    return out


def steady_state_noise(cfg, seed: int = 0) -> np.ndarray:
    """Z as [k, j] after getSteadyStateZ's burn-in."""
    if cfg.std_max < 0.0 or cfg.tau_stochastic <= 0.0:
        raise ValueError("the stochastic friction needs std_max >= 0 and tau_stochastic > 0")
    rng = np.random.default_rng(seed)
    tau = cfg.tau_stochastic
    dt = 0.1 * tau
    n_iter = int(math.ceil(10.0 * tau / dt))
    kern = (gaussian_kernel(cfg.length_spatial_corr, cfg.dx)
            if cfg.length_spatial_corr > cfg.dx else None)
    Z = np.zeros((cfg.ny, cfg.nx))
    for _ in range(n_iter):
        noise = rng.standard_normal((cfg.ny, cfg.nx))
        if kern is not None:
            noise = _convolve(noise, kern)
        Z = Z - (dt / tau) * Z + cfg.std_max * math.sqrt(2.0 * (dt / tau)) * noise
        if cfg.sym_noise > 0.0:
            Z = np.abs(Z) ** cfg.noise_pow_val
        elif cfg.sym_noise < 0.0:
            Z = -(np.abs(Z) ** cfg.noise_pow_val)
    return Z


def transport_noise(cfg, q: np.ndarray, Z: np.ndarray) -> np.ndarray:
    """q [k, j, var] with the transported noise set to q1 * Z, as
    getSteadyStateZ does after its burn-in when stoch_transport_flag is on
    (src/stochastic_module.f90:110-126: qp = Z and q = q1 * Z in the solved
    cells; elsewhere q1 is 0)."""
    if not cfg.stoch_transport_flag:
        raise ValueError("transport_noise: the run does not transport the noise")
    if q.shape[-1] != cfg.n_vars or Z.shape != q.shape[:2]:
        raise ValueError("transport_noise: q must be [ny, nx, n_vars] and Z [ny, nx]")
    out = np.array(q, dtype=np.float64, copy=True)
    out[..., 4 + cfg.n_solid] = out[..., 0] * Z
    return out
