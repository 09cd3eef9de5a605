"""GPU: accuracy of the reduced-precision paths on the 256x256x512
Rayleigh-Plateau workload itself (the box the MLUPS metric is quoted on):
periodic, cylinder of R=40 along z through (128.5, 128.5), 8 cosine modes
with amplitudes uniform in [0, 0.02 R] from seed 1234, tau_red=1.0,
tau_blue=0.7 (harmonic viscosity, S:674-680), sigma=A=0.01, beta=0.7, initial
state of the initialization form (S:850-896) generated on the device.

After 1000 steps the plain FP64 path (red/blue layout, which reproduces the
FP64 oracle bit for bit step by step: test_gpu_parity.py) is the reference:

* FP32 shifted storage in the factorised layout with the scatter-form phase
  field (the headline configuration): phi and u within 1e-5 L2-relative of
  FP64;
* FP16 shifted storage (FP32 phi and collision): phi within 5e-3 and u within
  5e-2 L2-relative, the bounds of the small-box tests;
* rho, phi and u errors (L2- and Linf-relative) of both are printed.

A reduced box with the same parameters (64x64x128, R=10, the same 8 modes
scaled by R/40) checks the scatter and the gather form of FP32 against FP64
after the same 1000 steps: phi within 1e-5 and u within 2e-5 (measured
1.0e-5 gather, 1.4e-5 scatter; FP32 storage with FP64 arithmetic reaches
3.3e-6 there, so the rest is the FP32 arithmetic of the step).

Where FP32 loses precision, and what the kernels do about it: the colour
split of a nearly pure node.  A colour that is nearly absent (rho_B ~ 1e-8
inside the red thread) formed as rho - rho_R, or as the complement of an FP32
majority term, carries noise of 1e-8..1e-7, above the 1e-8 evaporation limit
of Recolor (S:720-725); the recolouring then gathers that noise into blue
nuclei inside the red core, which the FP64 path never has.  Every shifted
path therefore forms the minority colour directly (sums, shifts, recolour
fractions, the scatter form's fixed-point terms and phase buffer), and these
tests pin the result against FP64 at the workload's own size.
"""
import os

import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

STEPS = 1000


def _print(*a):
    print(*a, flush=True)


def _run(cfg, precision, layout, steps, scatter=None, nx=None, ny=None, nz=None, r0=None, amps=None):
    nx, ny, nz = nx or cfg["nx"], ny or cfg["ny"], nz or cfg["nz"]
    saved = os.environ.get("CGL_SCATTER")
    if scatter is not None:
        os.environ["CGL_SCATTER"] = "1" if scatter else "0"
    try:
        lat = symmetry_amd.Lattice(nx, ny, nz, precision=precision, sigma=cfg["sigma"], rho0=cfg["rho0"],
                                   nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], layout=layout)
    finally:
        if saved is None:
            os.environ.pop("CGL_SCATTER", None)
        else:
            os.environ["CGL_SCATTER"] = saved
    lat.init_cylinder(nx / 2 + 0.5, ny / 2 + 0.5, r0 or cfg["r0"], 6.0 * cfg["beta"] * 0.134,
                      cfg["amps"] if amps is None else amps)
    lat.step(steps)
    m = lat.calmacros()
    form = lat.scatter_form()
    lat.close()
    u = np.stack([m["ux"], m["uy"], m["uz"]])
    return {"rho": m["rho"], "phi": m["phi"], "u": u}, form


def errors(m, ref, rho0=1.0):
    """L2- and Linf-relative errors of rho - rho0, phi and u against ref."""
    return {"rho": rel_err(m["rho"] - rho0, ref["rho"] - rho0), "phi": rel_err(m["phi"], ref["phi"]),
            "u": rel_err(m["u"], ref["u"])}


@pytest.fixture(scope="module")
def ref64():
    cfg = ic.rp_config3()
    _print("config3 f64-rb", cfg["nx"], cfg["ny"], cfg["nz"], STEPS, "steps")
    m, _ = _run(cfg, "f64", "rb", STEPS)
    assert np.all(np.isfinite(m["phi"])) and np.all(np.isfinite(m["u"]))
    # the flow has developed: the thread moves and phi spans both phases
    assert np.max(np.abs(m["u"])) > 1e-6 and m["phi"].min() < -0.99 and m["phi"].max() > 0.99
    return cfg, m


def test_config3_f32_scatter_within_1e5_of_f64(ref64):
    cfg, m64 = ref64
    _print("config3 f32-fs", STEPS, "steps")
    m, form = _run(cfg, "f32", "fs", STEPS)
    assert form, "the headline box runs the scatter-form phase field"
    e = errors(m, m64, cfg["rho0"])
    _print("config3 f32-fs scatter vs f64 (L2, Linf):", e)
    assert e["phi"][0] < 1e-5 and e["u"][0] < 1e-5


def test_config3_f16_storage_vs_f64(ref64):
    cfg, m64 = ref64
    _print("config3 f16-fs", STEPS, "steps")
    m, form = _run(cfg, "f16", "fs", STEPS)
    assert form
    e = errors(m, m64, cfg["rho0"])
    _print("config3 f16-fs scatter vs f64 (L2, Linf):", e)
    assert e["phi"][0] < 5e-3 and e["u"][0] < 5e-2


@pytest.mark.parametrize("scatter", [True, False])
def test_config3_parameters_reduced_box(scatter):
    """64x64x128 with tau_b=0.7 and the 8 modes scaled to R=10, 1000 steps."""
    cfg = ic.rp_config3()
    amps = tuple(a * 10.0 / 40.0 for a in cfg["amps"])
    shape = dict(nx=64, ny=64, nz=128, r0=10.0, amps=amps)
    _print("config3 parameters on 64x64x128, scatter", scatter)
    m64, _ = _run(cfg, "f64", "rb", STEPS, **shape)
    m32, form = _run(cfg, "f32", "fs", STEPS, scatter=scatter, **shape)
    assert form == scatter
    e = errors(m32, m64, cfg["rho0"])
    _print("64x64x128 f32-fs", "scatter" if scatter else "gather", "vs f64 (L2, Linf):", e)
    assert e["phi"][0] < 1e-5 and e["u"][0] < 2e-5
