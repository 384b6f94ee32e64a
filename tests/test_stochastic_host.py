"""The host set-up of the stochastic friction (imex_sflow2d_dev_amd/
stochastic.py) against what the mathematics of getSteadyStateZ fixes
(src/stochastic_module.f90:79-132): the random numbers are numpy's, so the
sample is pinned through its distribution, not value by value."""
import math

import numpy as np
import pytest

from imex_sflow2d_dev_amd import stochastic
from imex_sflow2d_dev_amd.config import RunConfig


def _cfg(**kw):
    c = RunConfig(nx=200, ny=150, dx=10.0)
    c.stochastic_flag = True
    c.std_max, c.tau_stochastic = 3.0, 2.0
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def test_burn_in_reaches_the_euler_maruyama_stationary_variance():
    """Z_{n+1} = (1 - a) Z_n + sigma sqrt(2a) N with a = dt / tau = 0.1 from
    Z = 0 over 100 steps: mean 0, variance sigma^2 (1 - (1-a)^200) / (1 - a/2)."""
    c = _cfg()
    Z = stochastic.steady_state_noise(c, seed=1)
    a = 0.1
    var = c.std_max ** 2 * (1.0 - (1.0 - a) ** 200) / (1.0 - a / 2.0)
    n = Z.size
    assert abs(Z.mean()) < 5.0 * math.sqrt(var / n)
    # the sample variance of n normal values: relative standard error sqrt(2/n)
    assert abs(Z.var() / var - 1.0) < 5.0 * math.sqrt(2.0 / n)
    # no spatial correlation without the kernel
    r = np.corrcoef(Z[:, :-1].ravel(), Z[:, 1:].ravel())[0, 1]
    assert abs(r) < 5.0 / math.sqrt(n)


def test_kernel_and_spatial_correlation():
    k = stochastic.gaussian_kernel(45.0, 10.0)
    assert k.shape == (5, 5)
    assert abs(k.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(k, k.T, rtol=0, atol=0)
    np.testing.assert_allclose(k, k[::-1, ::-1], rtol=1e-15)
    assert k[2, 2] == k.max()
    # a unit impulse comes back as the kernel, centred on it
    imp = np.zeros((11, 11))
    imp[5, 5] = 1.0
    np.testing.assert_allclose(stochastic._convolve(imp, k)[3:8, 3:8], k[::-1, ::-1], rtol=1e-15)
    c = _cfg(length_spatial_corr=45.0)
    Z = stochastic.steady_state_noise(c, seed=2)
    r = np.corrcoef(Z[:, :-1].ravel(), Z[:, 1:].ravel())[0, 1]
    assert r > 0.5                   # neighbours now share most of their noise


@pytest.mark.parametrize("sym", [1.0, -1.0])
def test_sym_noise_map(sym):
    c = _cfg(sym_noise=sym, noise_pow_val=0.5)
    Z = stochastic.steady_state_noise(c, seed=3)
    assert (np.sign(Z) == sym).mean() > 0.99 and ((Z == 0) | (np.sign(Z) == sym)).all()


def test_requires_parameters():
    with pytest.raises(ValueError):
        stochastic.steady_state_noise(_cfg(tau_stochastic=-1.0))
