"""Linear capillary instability of a liquid thread in another liquid: the
growth rate of an axisymmetric mode, for the Rayleigh-Plateau checks of the
GPU runs.

A thread of radius R (fluid 1: density rho1, kinematic viscosity nu1) in an
unbounded fluid 2 (rho2, nu2), surface tension sigma, perturbed by
eta exp(ikz + omega t).  Linearised Navier-Stokes in each fluid: a potential
part (pressure -rho omega Phi, Phi = A I0(kr) inside, D K0(kr) outside) plus
a vortical part (u_r = i k C I1(l r), u_z = -C l I0(l r) inside; K1, K0
outside) with l^2 = k^2 + omega / nu.  At r = R: u_r and u_z continuous,
tangential stress continuous, and the normal-stress jump balancing
sigma (k^2 - 1/R^2) eta with omega eta = u_r.  The four conditions are a
4 x 4 linear system in the four amplitudes; omega is the positive real root
of its determinant.  Exponentially scaled Bessel functions keep the columns
finite at small viscosity (scaling a column leaves the roots unchanged).

Closed-form limits, checked on the CPU (tests/test_linear_theory.py):
viscosity to zero gives the inviscid two-fluid relation
omega^2 = sigma x (1 - x^2) / (R^3 (rho1 I0/I1 + rho2 K0/K1)), x = kR, which
for rho2 = 0 is Rayleigh's; a viscous thread in a negligible outer fluid at
long wavelength and negligible inertia gives Trouton's
omega = sigma (1 - x^2) / (6 mu1 R).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.optimize import brentq
from scipy.special import ive, kve


def _det(omega, k, R, sigma, rho1, rho2, nu1, nu2):
    mu1, mu2 = rho1 * nu1, rho2 * nu2
    l1 = math.sqrt(k * k + omega / nu1)
    l2 = math.sqrt(k * k + omega / nu2)
    a, b1, b2 = k * R, l1 * R, l2 * R
    # column 1, 3: potential parts (I(a) scaled by e^-a, K(a) by e^a);
    # column 2: inner vortical (I(b1) e^-b1); column 4: outer vortical (K(b2) e^b2)
    I0a, I1a = ive(0, a), ive(1, a)
    K0a, K1a = kve(0, a), kve(1, a)
    I0b, I1b = ive(0, b1), ive(1, b1)
    K0b, K1b = kve(0, b2), kve(1, b2)
    dI1a = I0a - I1a / a  # I1'(a), same scaling
    dI1b = I0b - I1b / b1
    dK1a = -K0a - K1a / a
    dK1b = -K0b - K1b / b2
    s = sigma * (1.0 - a * a) / (R * R)  # normal-stress row multiplied by omega
    M = np.array([
        [I1a, I1b, K1a, -K1b],
        [k * I0a, l1 * I0b, -k * K0a, l2 * K0b],
        [2 * k * k * mu1 * I1a, mu1 * (k * k + l1 * l1) * I1b, 2 * k * k * mu2 * K1a, -mu2 * (k * k + l2 * l2) * K1b],
        [-rho1 * omega * omega * I0a - 2 * mu1 * omega * k * k * dI1a + s * k * I1a,
         -2 * mu1 * omega * k * l1 * dI1b + s * k * I1b,
         rho2 * omega * omega * K0a - 2 * mu2 * omega * k * k * dK1a,
         2 * mu2 * omega * k * l2 * dK1b],
    ])
    return np.linalg.det(M)


def inviscid_rate(k, R, sigma, rho1, rho2):
    """Growth rate of the inviscid two-fluid relation (0 for kR >= 1)."""
    from scipy.special import iv, kv
    x = k * R
    if x >= 1.0:
        return 0.0
    den = rho1 * iv(0, x) / iv(1, x) + rho2 * kv(0, x) / kv(1, x)
    return math.sqrt(sigma * x * (1.0 - x * x) / (R ** 3 * den))


def growth_rate(k, R, sigma, rho1, rho2, nu1, nu2):
    """The positive real growth rate of the axisymmetric mode k (0 if stable).

    Viscosity only slows the instability, so the root lies below the
    inviscid rate (plus the inertia-free viscous bound when the fluids are
    very light); it is bracketed by a logarithmic scan and refined by Brent's
    method."""
    if k * R >= 1.0:
        return 0.0
    hi = inviscid_rate(k, R, sigma, rho1, rho2) * 1.001
    f = lambda w: _det(w, k, R, sigma, rho1, rho2, nu1, nu2)
    grid = hi * np.logspace(-8, 0, 400)
    vals = [f(w) for w in grid]
    for i in range(len(grid) - 1, 0, -1):  # the largest root
        if vals[i] == 0.0:
            return float(grid[i])
        if np.sign(vals[i]) != np.sign(vals[i - 1]):
            return float(brentq(f, grid[i - 1], grid[i], xtol=1e-14 * hi, rtol=1e-12))
    raise ValueError("no growing root found")
