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
"""
import math

import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic

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
