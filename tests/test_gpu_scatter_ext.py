"""The scatter form with wall faces and moving particles (FP32 and FP16
shifted storage, factorised layout): the next phase field summed in fixed
point from the terms every node hands downstream, with the terms that leave
through a wall returned to the node (half-way bounce-back, S:1557-1590), the
particles' band updated by a per-node kernel that adds its terms atomically,
and the links' bounce-back values (S:1004-1018) added by the curve kernel.

Pinned against the plain FP64 path (which the oracle pins bit for bit,
test_gpu_ext) and against the gather form of the same precision (the
moments pass per step): the scatter form must be at least as close to FP64
as the gather form is, within the same tolerances as the gather form's
tests.  Sizes are the particle workload's 40^3 stand-in (one sphere at a flat
interface between z walls) and two spheres whose band boxes overlap, so
that a node in two boxes adds its terms once.
"""
import os

import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import rel_err
from tests.test_gpu_ext import _lattice, _lattice_particles, _live, _particle_setup, _wall_cfg

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _with_scatter(on, make):
    """A lattice made with the scatter form allowed (on) or refused."""
    old = os.environ.get("CGL_SCATTER")
    os.environ["CGL_SCATTER"] = "1" if on else "0"
    try:
        return make()
    finally:
        if old is None:
            del os.environ["CGL_SCATTER"]
        else:
            os.environ["CGL_SCATTER"] = old


def test_scatter_form_taken_with_walls_and_particles():
    """Shifted storage in the factorised layout takes the scatter form with
    wall faces and particles; FP64, the red/blue layout and mirror faces keep
    the moments pass."""
    cfg, parts, ri, bi = _particle_setup()
    for prec, layout, want in (("f32", "fs", True), ("f16", "fs", True), ("f64", "fs", False), ("f32", "rb", False)):
        lat = _lattice_particles(cfg, parts, ri, bi, prec, layout=layout)
        assert lat.scatter_form() == want, (prec, layout)
        lat.close()
    mcfg = _wall_cfg(18, 16, 26, (2, 2, 0, 0, 1, 1))
    lat = _lattice(mcfg, "f32", layout="fs")
    assert not lat.scatter_form()
    lat.close()
    wcfg = _wall_cfg(32, 16, 24, (0, 0, 0, 0, 1, 1))
    lat = _lattice(wcfg, "f32", layout="fs")
    assert lat.scatter_form()
    lat.close()


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_walls_scatter_close_to_f64(prec):
    """A perturbed flat interface between z walls, 200 steps (the geometry of
    test_walls_f32_close_to_f64): phi and u of the scatter form against FP64,
    and no farther from it than the gather form of the same precision, with
    a margin for the rounding of a different summation: 2 for FP32, 4 for
    FP16 storage, whose half-precision rounding of the populations differs
    between the two forms (measured 2.6 for phi, both 10 times below the
    tolerance)."""
    cfg = _wall_cfg(32, 16, 24, (0, 0, 0, 0, 1, 1))
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=12.5, amp=2.0, rho0=1.0)
    live = _live(ri.shape[:3], cfg["bc"])
    out = {}
    for key, p, layout, sc in (("f64", "f64", "rb", False), ("gather", prec, "fs", False), ("scatter", prec, "fs", True)):
        lat = _with_scatter(sc, lambda: _lattice(cfg, p, layout=layout))
        assert lat.scatter_form() == sc
        lat.upload(ri, bi)
        lat.step(200)
        out[key] = lat.calmacros()
        lat.close()
    err = {}
    for key in ("gather", "scatter"):
        phi, _ = rel_err(out[key]["phi"][live], out["f64"]["phi"][live])
        u = np.stack([out[key][k][live] for k in ("ux", "uy", "uz")])
        u64 = np.stack([out["f64"][k][live] for k in ("ux", "uy", "uz")])
        err[key] = (phi, rel_err(u, u64)[0])
    _print(prec, "phi / u rel err: gather", err["gather"], "scatter", err["scatter"])
    tol_phi, tol_u = (1e-5, 3e-5) if prec == "f32" else (2e-5, 5e-4)
    assert err["scatter"][0] < tol_phi and err["scatter"][1] < tol_u
    margin = 2 if prec == "f32" else 4
    assert err["scatter"][0] < margin * err["gather"][0] + 1e-7
    assert err["scatter"][1] < margin * err["gather"][1] + 1e-7


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_particle_scatter_close_to_f64(prec):
    """A moving sphere at the interface, 30 steps: masks exact, fluid phi
    and the particle's centre against FP64 (FP32 within the gather form's
    tolerances; FP16 storage within 1e-5 for phi and 1e-4 for the centre)."""
    cfg, parts, ri, bi = _particle_setup()
    res = {}
    for key, p, layout in (("f64", "f64", "rb"), ("scatter", prec, "fs")):
        lat = _lattice_particles(cfg, parts, ri, bi, p, layout=layout)
        assert lat.scatter_form() == (key == "scatter")
        lat.step(30)
        res[key] = (lat.calmacros(), lat.get_particles(), lat.masks())
        lat.close()
    assert all(np.array_equal(res["scatter"][2][k], res["f64"][2][k]) for k in res["f64"][2])
    fluid = res["f64"][2]["walls_all"] == 0
    l2, linf = rel_err(res["scatter"][0]["phi"][fluid], res["f64"][0]["phi"][fluid])
    dc = np.max(np.abs(res["scatter"][1]["center"] - res["f64"][1]["center"]))
    _print(prec, "phi", l2, linf, "centre", dc)
    assert l2 < 1e-5
    assert dc < (1e-5 if prec == "f32" else 1e-4)


def test_observation_between_scatter_runs():
    """Reading the fields between runs (calmacros writes the phase-field
    buffer as phi) makes the next step take the current phase field from a
    moments pass; the trajectory continues within rounding of one
    uninterrupted run."""
    cfg, parts, ri, bi = _particle_setup()
    a = _lattice_particles(cfg, parts, ri, bi, "f32", layout="fs")
    b = _lattice_particles(cfg, parts, ri, bi, "f32", layout="fs")
    a.step(10)
    a.calmacros()
    a.step(20)
    b.step(30)
    ma, mb = a.calmacros(), b.calmacros()
    fluid = b.masks()["walls_all"] == 0
    l2, _ = rel_err(ma["phi"][fluid], mb["phi"][fluid])
    dc = np.max(np.abs(a.get_particles()["center"] - b.get_particles()["center"]))
    _print("observed vs uninterrupted: phi", l2, "centre", dc)
    assert l2 < 1e-6 and dc < 1e-6


def test_overlapping_bands_scatter_close_to_f64():
    """Two spheres 2R + 5 apart at the interface (their band boxes, R + 3
    around each centre, overlap): a node in both boxes adds its terms once,
    so the FP32 scatter form stays within the gather form's tolerance of
    FP64 and the total density of the fluid is conserved as in FP64."""
    n, rad = 40, 6.0
    cfg = ic.particle_config2(n=n, rad=rad)
    c0 = np.array(cfg["center"]) + np.array([0.3, -0.2, 0.4])
    centers = [c0 - np.array([rad + 2.5, 0.0, 0.0]), c0 + np.array([rad + 2.5, 0.0, 0.0])]
    parts = ic.sphere_particles(centers, rad, cfg["rho_p"], cfg["cita_deg"], cfg["sigma"])
    parts["up"] = np.array([[0.02, 0.01, -0.01], [-0.02, 0.0, 0.01]], dtype=np.float64)
    parts["rp"] = np.zeros((2, 3), dtype=np.float64)
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    res = {}
    for prec, layout in (("f64", "rb"), ("f32", "fs")):
        lat = _lattice_particles(cfg, parts, ri, bi, prec, layout=layout)
        assert lat.scatter_form() == (prec == "f32")
        lat.step(40)
        res[prec] = (lat.calmacros(), lat.get_particles(), lat.masks())
        lat.close()
    assert all(np.array_equal(res["f32"][2][k], res["f64"][2][k]) for k in res["f64"][2])
    fluid = res["f64"][2]["walls_all"] == 0
    l2, _ = rel_err(res["f32"][0]["phi"][fluid], res["f64"][0]["phi"][fluid])
    mass = [float(np.sum(res[k][0]["rho"][fluid])) for k in ("f32", "f64")]
    dc = np.max(np.abs(res["f32"][1]["center"] - res["f64"][1]["center"]))
    _print("two spheres: phi", l2, "mass", mass, "centre", dc)
    assert l2 < 1e-5 and dc < 1e-5
    assert abs(mass[0] - mass[1]) < 1e-6 * mass[1]


@pytest.mark.parametrize("prec", ["f32", "f16"])
@pytest.mark.parametrize("case", ["x", "xyz"])
def test_wall_drop_scatter_close_to_f64(case, prec):
    """A drop touching walls in x only, and in a closed box (walls on all six
    faces, so edges and corners where two or three walls return a term), 100
    steps: phi of the scatter form against FP64 and no farther from it than
    the gather form of the same precision, with the margins of
    test_walls_scatter_close_to_f64."""
    from tests.test_gpu_ext import WALL_CASES
    cfg, drop = WALL_CASES[case]
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=-1e3, rho0=1.0, drop=drop,
                                        sigma=cfg["sigma"])
    live = _live(ri.shape[:3], cfg["bc"])
    out = {}
    for key, p, layout, sc in (("f64", "f64", "rb", False), ("gather", prec, "fs", False), ("scatter", prec, "fs", True)):
        lat = _with_scatter(sc, lambda: _lattice(cfg, p, layout=layout))
        assert lat.scatter_form() == sc
        lat.upload(ri, bi)
        lat.step(100)
        out[key] = lat.calmacros()
        lat.close()
    err = {k: rel_err(out[k]["phi"][live], out["f64"]["phi"][live])[0] for k in ("gather", "scatter")}
    _print(case, prec, "phi rel err: gather", err["gather"], "scatter", err["scatter"])
    assert err["scatter"] < (1e-5 if prec == "f32" else 2e-5)
    margin = 2 if prec == "f32" else 4
    assert err["scatter"] < margin * err["gather"] + 1e-7


@pytest.mark.parametrize("prec", ["f32", "f16"])
def test_launch_shape_bit_identical_with_particles(prec):
    """X chunks 8, 2 and 6 (the last leaves a short chunk at the box's end)
    give bit-identical states and particle centres with z walls and two
    spheres whose band boxes overlap: every term is an integer in fixed
    point, so the launch shape changes no sum."""
    n, rad = 40, 6.0
    cfg = ic.particle_config2(n=n, rad=rad)
    c0 = np.array(cfg["center"]) + np.array([0.3, -0.2, 0.4])
    centers = [c0 - np.array([rad + 2.5, 0.0, 0.0]), c0 + np.array([rad + 2.5, 0.0, 0.0])]
    parts = ic.sphere_particles(centers, rad, cfg["rho_p"], cfg["cita_deg"], cfg["sigma"])
    parts["up"] = np.array([[0.02, 0.01, -0.01], [-0.02, 0.0, 0.01]], dtype=np.float64)
    parts["rp"] = np.zeros((2, 3), dtype=np.float64)
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    res = []
    for xc in (8, 2, 6):
        lat = _lattice(cfg, prec, npar=2, layout="fs", xchunk=xc)
        lat.upload(ri, bi)
        lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"],
                          parts["fg"], up=parts["up"], rp=parts["rp"])
        assert lat.scatter_form()
        lat.step(25)
        tp = lat.tuned_params()
        assert tp["xchunk"] == xc
        res.append((lat.download(), lat.get_particles()["center"]))
        lat.close()
    for (r, b), c in res[1:]:
        assert np.array_equal(r, res[0][0][0]) and np.array_equal(b, res[0][0][1])
        assert np.array_equal(c, res[0][1])
