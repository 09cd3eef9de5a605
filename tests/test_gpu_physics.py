"""Physics of the Rayleigh-Plateau workload (32x32x128 periodic, R=8, 10 %
mode, tau=1, sigma=0.01): the FP32 shifted-storage path against the FP64
path, which reproduces the FP64 oracle bit for bit step by step
(test_gpu_parity.py), through the perturbation's growth and the breakup of
the thread.

* Neck and crest radii R(z) = sqrt(red area / pi) of every z-slice agree to
  1e-3 relative until the neck is two lattice units thick.
* The exponential growth rate of the amplitude (R_max - R_min)/2, fitted over
  steps 2000-8000, agrees to 0.2 %.
* Breakup (the first check where phi on the axis at the neck turns blue)
  happens in the same 100-step window.

and, in wide cross-sections, the fitted growth rate of a mode against the
linear dispersion relation of a viscous thread in a viscous fluid, and the
Laplace pressure jump of cylinders and a sphere.
"""
import math

import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests import linear_theory as lt

pytestmark = pytest.mark.gpu

CHECK = 100
MAX_STEPS = 30000


def _run(precision, layout):
    cfg = ic.rp_config1()
    ri, bi, _ = ic.workload_state(cfg, exact=False)
    lat = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                               rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], layout=layout)
    lat.upload(ri, bi)
    cx, cy = cfg["nx"] // 2, cfg["ny"] // 2
    t, rmin, rmax, breakup = [], [], [], None
    for k in range(MAX_STEPS // CHECK + 1):
        m = lat.calmacros()
        red = ((1.0 + m["phi"]) / 2.0).sum(axis=(0, 1))
        r = np.sqrt(red / math.pi)
        t.append(k * CHECK)
        rmin.append(r.min())
        rmax.append(r.max())
        if m["phi"][cx - 1:cx + 1, cy - 1:cy + 1, :].mean(axis=(0, 1)).min() < 0.0:
            breakup = k * CHECK
            break
        lat.step(CHECK)
    lat.close()
    return np.array(t), np.array(rmin), np.array(rmax), breakup


def _rate(t, amp, t0=2000, t1=8000):
    sel = (t >= t0) & (t <= t1)
    return np.polyfit(t[sel], np.log(amp[sel]), 1)[0]


def test_rayleigh_plateau_growth_and_breakup_f32_vs_f64():
    t64, lo64, hi64, b64 = _run("f64", "rb")
    t32, lo32, hi32, b32 = _run("f32", "fs")
    print("breakup step f64", b64, "f32", b32, flush=True)
    assert b64 is not None and b32 is not None, "no breakup within %d steps" % MAX_STEPS
    assert b32 == b64
    n = min(len(t64), len(t32))
    thick = lo64[:n] > 2.0
    assert np.max(np.abs(lo32[:n][thick] / lo64[:n][thick] - 1.0)) < 1e-3
    assert np.max(np.abs(hi32[:n] / hi64[:n] - 1.0)) < 1e-3
    w64 = _rate(t64, (hi64 - lo64) / 2.0)
    w32 = _rate(t32, (hi32 - lo32) / 2.0)
    print("growth rate per step f64 %.6e f32 %.6e" % (w64, w32), flush=True)
    assert w64 > 0.0
    assert abs(w32 / w64 - 1.0) < 2e-3


KPROF = 6.0 * 0.7 * 0.134


def _mode_growth(n, nz, r0, tau, a0, every=100, max_steps=40000):
    """FP64 run of a z-periodic thread of radius r0 with one cosine mode of
    amplitude a0 and wavelength nz in an n x n cross-section; the mode's
    amplitude from the red area of each z slice, R(z) = sqrt(area / pi),
    projected on cos and sin of k z, every ``every`` steps until it reaches
    r0 / 4."""
    nu = (tau - 0.5) / 3.0
    lat = symmetry_amd.Lattice(n, n, nz, precision="f64", sigma=0.01, rho0=1.0, nu_r=nu, nu_b=nu, beta=0.7)
    lat.init_cylinder(n / 2 + 0.5, n / 2 + 0.5, r0, KPROF, (a0,))
    kz = 2.0 * math.pi * np.arange(1, nz + 1) / nz
    t, amp = [], []
    for k in range(max_steps // every + 1):
        r = np.sqrt(lat.diagnostics(slices=True)["slice_red"] / math.pi)
        t.append(k * every)
        amp.append(2.0 / nz * math.hypot(float(np.dot(r, np.cos(kz))), float(np.dot(r, np.sin(kz)))))
        if amp[-1] > 0.25 * r0:
            break
        lat.step(every)
    lat.close()
    return np.array(t), np.array(amp)


@pytest.mark.parametrize("n,nz,r0,tau", [(128, 128, 8.0, 1.0), (144, 126, 12.0, 0.7), (144, 126, 12.0, 1.0)])
def test_growth_rate_against_linear_theory(n, nz, r0, tau):
    """The fitted exponential growth rate of the mode against the positive
    root of the linear dispersion relation of a viscous thread in an
    unbounded viscous fluid of the same density and viscosity
    (tests/linear_theory.py), while the amplitude grows from 2 a0 to
    0.15 r0.  The cross-section is wide enough (about 3/k or more from the
    axis to the box face) for the outer flow of the unbounded theory.  The
    first case is the 32x32x128 workload's mode (R = 8, wavelength 128,
    tau = 1), whose own 32 x 32 cross-section confines the outer return
    flow (its rate there, 9.19e-5 per step, is 0.45 of this one).  Measured
    ratios to the theory: 0.983, 0.984 and 0.987."""
    a0 = 0.01 * r0
    t, amp = _mode_growth(n, nz, r0, tau, a0)
    nu = (tau - 0.5) / 3.0
    theory = lt.growth_rate(2.0 * math.pi / nz, r0, 0.01, 1.0, 1.0, nu, nu)
    sel = (amp >= 2.0 * a0) & (amp <= 0.15 * r0)
    assert np.count_nonzero(sel) >= 10, (t, amp)
    rate = np.polyfit(t[sel], np.log(amp[sel]), 1)[0]
    print("R %g x %.4f tau %g: fitted %.5e per step over steps %d-%d, linear theory %.5e, ratio %.4f"
          % (r0, 2 * math.pi * r0 / nz, tau, rate, t[sel][0], t[sel][-1], theory, rate / theory), flush=True)
    assert abs(rate / theory - 1.0) < 0.03


def _jump(lat, inside, outside):
    m = lat.calmacros()
    return float(m["rho"][inside].mean() - m["rho"][outside].mean())


@pytest.mark.parametrize("r0", [10.0, 14.0, 18.0])
def test_laplace_law_cylinder(r0):
    """A red cylinder along z in a periodic box, set up at one density in
    both fluids (no pressure jump), relaxes to the Laplace jump of the CSF
    surface tension: rho_in - rho_out = 3 sigma / R (p = rho / 3, kappa =
    1 / R).  The density is averaged more than 4 nodes inside the interface
    and more than 6 outside it.  Measured ratios to 3 sigma / R: 1.0155,
    1.0073, 1.0038 (the diffuse interface's share shrinks with R)."""
    n, nz, sigma = 64, 8, 0.01
    cfg = dict(nx=n, ny=n, nz=nz, r0=r0, amps=(), sigma=0.0, rho0=1.0, nu_r=1.0 / 6.0, nu_b=1.0 / 6.0, beta=0.7)
    ri, bi, _ = ic.workload_state(cfg, exact=False)
    lat = symmetry_amd.Lattice(n, n, nz, precision="f64", sigma=sigma, rho0=1.0, nu_r=1.0 / 6.0, nu_b=1.0 / 6.0,
                               beta=0.7)
    lat.upload(ri, bi)
    x = np.arange(1, n + 1) - (n / 2 + 0.5)
    r = np.sqrt(x[:, None] ** 2 + x[None, :] ** 2)[:, :, None] * np.ones((1, 1, nz))
    inside, outside = r < r0 - 4.0, r > r0 + 6.0
    assert abs(_jump(lat, inside, outside)) < 1e-12
    lat.step(4000)
    d = _jump(lat, inside, outside)
    print("cylinder R %g: jump %.6e, 3 sigma / R %.6e, ratio %.4f" % (r0, d, 3 * sigma / r0, d * r0 / (3 * sigma)),
          flush=True)
    assert abs(d * r0 / (3 * sigma) - 1.0) < 0.03


def test_laplace_law_sphere():
    """A red sphere of radius 12 in a periodic 64^3 box, set up without a
    pressure jump, relaxes to rho_in - rho_out = 6 sigma / R (kappa = 2 / R);
    measured ratio 1.0074."""
    n, r0, sigma = 64, 12.0, 0.01
    c = n / 2 + 0.5
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=-1e3, rho0=1.0, beta=0.7, drop=(c, c, c, r0), sigma=0.0)
    lat = symmetry_amd.Lattice(n, n, n, precision="f64", sigma=sigma, rho0=1.0, nu_r=1.0 / 6.0, nu_b=1.0 / 6.0,
                               beta=0.7)
    lat.upload(ri, bi)
    x = np.arange(1, n + 1) - c
    r = np.sqrt(x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    inside, outside = r < r0 - 4.0, r > r0 + 6.0
    assert abs(_jump(lat, inside, outside)) < 1e-12
    lat.step(4000)
    d = _jump(lat, inside, outside)
    print("sphere R %g: jump %.6e, 6 sigma / R %.6e, ratio %.4f" % (r0, d, 6 * sigma / r0, d * r0 / (6 * sigma)),
          flush=True)
    assert abs(d * r0 / (6 * sigma) - 1.0) < 0.03
