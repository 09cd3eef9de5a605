"""The linear dispersion relation of a viscous thread in a viscous fluid
(tests/linear_theory.py), which the GPU growth-rate test compares against,
pinned by its closed-form limits."""
import math

import pytest

from tests import linear_theory as lt
from tests.helpers import rp_growth_rate


@pytest.mark.parametrize("x", [0.2, 0.39, 0.6, 0.8])
def test_inviscid_limit(x):
    """Viscosity to zero: the inviscid two-fluid relation; with a negligible
    outer fluid, Rayleigh's omega^2 = sigma/(rho R^3) x I1/I0 (1 - x^2)."""
    R, sigma = 10.0, 0.01
    k = x / R
    w = lt.growth_rate(k, R, sigma, 1.0, 1.0, 1e-10, 1e-10)
    assert abs(w / lt.inviscid_rate(k, R, sigma, 1.0, 1.0) - 1.0) < 1e-3
    w1 = lt.growth_rate(k, R, sigma, 1.0, 1e-12, 1e-10, 1e-10)
    assert abs(w1 / math.sqrt(sigma / R ** 3 * rp_growth_rate(x)) - 1.0) < 1e-3
    # the outer fluid's inertia slows the thread
    assert w < w1


@pytest.mark.parametrize("x", [0.02, 0.05, 0.1])
def test_viscous_long_wave_limit(x):
    """A very viscous light thread in a negligible outer fluid at long
    wavelength: Trouton's omega = sigma (1 - x^2) / (6 mu R)."""
    R, sigma, mu, rho = 8.0, 0.01, 1.0, 1e-6
    w = lt.growth_rate(x / R, R, sigma, rho, 1e-12, mu / rho, 1.0)
    assert abs(w / (sigma * (1.0 - x * x) / (6.0 * mu * R)) - 1.0) < 1e-4


def test_viscosity_slows_and_short_waves_are_stable():
    R, sigma, k = 8.0, 0.01, 0.39 / 8.0
    rates = [lt.growth_rate(k, R, sigma, 1.0, 1.0, nu, nu) for nu in (1e-4, 0.01, 0.05, 1.0 / 6.0)]
    assert all(a > b for a, b in zip(rates, rates[1:]))
    assert lt.growth_rate(1.05 / R, R, sigma, 1.0, 1.0, 0.1, 0.1) == 0.0
