"""Shared test helpers: identical seeded inputs for the GPU library and the
oracle, and field comparisons."""
import math

import numpy as np

import oracle
from symmetry_amd import ic


def small_rp(nx=32, ny=32, nz=128, r0=8.0, amps=(0.8,), sigma=0.01, rho0=1.0, nu_r=1.0 / 6.0,
             nu_b=1.0 / 6.0, beta=0.7):
    cfg = dict(nx=nx, ny=ny, nz=nz, r0=r0, amps=amps, sigma=sigma, rho0=rho0, nu_r=nu_r, nu_b=nu_b, beta=beta)
    ri, bi, phi = ic.workload_state(cfg, exact=False)
    return cfg, ri, bi


def oracle_from(cfg, ri, bi):
    st = oracle.OracleState(cfg["nx"], cfg["ny"], cfg["nz"], bc=(0,) * 6, npar=0, sigma=cfg["sigma"],
                            rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"])
    st.ri[...] = ri
    st.bi[...] = bi
    return st


def rel_err(a, b):
    """L2-relative and Linf-relative difference of a against reference b."""
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    l2 = np.linalg.norm(d) / nb if nb > 0 else np.linalg.norm(d)
    mb = np.max(np.abs(b))
    linf = np.max(np.abs(d)) / mb if mb > 0 else np.max(np.abs(d))
    return l2, linf


def rp_growth_rate(kR):
    """Rayleigh's inviscid growth-rate factor  x I1(x)/I0(x) (1 - x^2), x = kR;
    omega^2 = sigma/(rho R^3) times this factor (single fluid)."""
    from scipy.special import iv
    return kR * iv(1, kR) / iv(0, kR) * (1.0 - kR * kR)


def mass(ri, bi):
    return float(np.sum(ri, dtype=np.float64)), float(np.sum(bi, dtype=np.float64))


__all__ = ["small_rp", "oracle_from", "rel_err", "mass", "math"]
