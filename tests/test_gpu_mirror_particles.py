"""GPU parity: moving particles in the reference's own face arrangement --
walls at x = 1 and z = 1, mirrors at x = xl, y = 1, y = yl and z = zl
(S:806-807, S:1443-1592) -- with particles centred on the mirror planes, the
reference's repulsion planes (S:1607-1612), Lagrange in-domain box (S:1832)
and symmetry-plane rules of movement (S:1272-1335), against the FP64
oracle, whose particle stages are pinned bit for bit to the compiled
reference with its 12 particles (tests/test_oracle_vs_ref.py).

Masks must match exactly (set_walls with its mirror-face copies, S:326-395);
fields and particle states within the tolerance of the other particle tests
(block-tree force sums on the GPU, serial sums in the oracle).
"""
import math

import numpy as np
import pytest

import oracle
import symmetry_amd
from symmetry_amd import ic

pytestmark = pytest.mark.gpu

MACROS = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")
W, M = symmetry_amd.BC_WALL, symmetry_amd.BC_MIRROR
REF_BC = (W, M, M, M, W, M)


def _print(*a):
    print(*a, flush=True)


def _setup(n=36, rad=5.0):
    """Three spheres: on the x = xl - 1/2 mirror plane, on the y = 3/2 mirror
    plane (particle par/3 + 1, whose x and z components the rules average) and
    one inside; velocities within their planes, so that the rules keep them
    there, and large enough to cover and uncover nodes in a few steps."""
    cfg = dict(nx=n, ny=n, nz=n, bc=REF_BC, sigma=0.005, rho0=1.0, nu_r=0.1, nu_b=0.1, beta=0.7, z0=n / 2.0)
    centers = [(n - 0.5, 18.3, 18.4), (15.2, 1.5, 20.3), (17.6, 23.1, 16.7)]
    parts = ic.sphere_particles(centers, rad, 1.5, 90.0, cfg["sigma"])
    parts["up"] = np.array([(0.0, 0.25, -0.05), (0.2, 0.0, 0.1), (0.3, -0.2, 0.05)])
    parts["rp"] = np.array([(0.02, 0.0, 0.0), (0.0, 0.02, 0.0), (0.01, 0.0, 0.02)])
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    # the reference's geometry for its box (S:1607-1612, S:1832)
    geo = dict(rep_plane=(1.5, n - 1.5, 1.5, n - 1.5, 1.5, n - 0.5), rep_on=(1,) * 6,
               dom_lo=(0.0, 1.5, 0.0), dom_hi=(n - 0.5, n - 0.5, n - 0.5),
               dom_on=(0, 0, 1, 1, 0, 1, 0, 0, 1), sym_rules=1)
    return cfg, parts, ri, bi, geo


def _oracle(cfg, parts, ri, bi, geo):
    npar = len(parts["rad"])
    st = oracle.OracleState(cfg["nx"], cfg["ny"], cfg["nz"], bc=cfg["bc"], npar=npar, sigma=cfg["sigma"],
                            rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], fg=parts["fg"])
    st.ri[...] = ri
    st.bi[...] = bi
    st.center[:npar] = parts["center"]
    st.center_old[:npar] = parts["center"]
    st.center0[:npar] = parts["center"]
    for k in ("rad", "mass", "inertia", "cita"):
        st.arrays[k][:npar] = parts[k]
    st.up[:npar] = parts["up"]
    st.rp[:npar] = parts["rp"]
    s = st.s
    for f in range(6):
        s.rep_plane[f] = geo["rep_plane"][f]
        s.rep_on[f] = geo["rep_on"][f]
    for a in range(3):
        s.dom_lo[a], s.dom_hi[a] = geo["dom_lo"][a], geo["dom_hi"][a]
        s.dom_lo_on[a], s.dom_lo_strict[a], s.dom_hi_on[a] = geo["dom_on"][3 * a:3 * a + 3]
    s.ref_rules = geo["sym_rules"]
    st.call("set_walls")
    st.ffi[...] = st.ri + st.bi  # initialization's ffi = fi (S:894)
    st.call("calmacros")
    st.call("derivatives")
    return st


def _lattice(cfg, parts, ri, bi, geo, precision="f64", layout="rb"):
    lat = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                               rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"],
                               bc=cfg["bc"], npar=len(parts["rad"]), layout=layout)
    lat.upload(ri, bi)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"],
                      up=parts["up"], rp=parts["rp"])
    lat.set_particle_geometry(**geo)
    return lat


def _live(shape):
    return tuple(slice(1, n - 1) for n in shape)


def test_mirror_masks_at_start():
    cfg, parts, ri, bi, geo = _setup()
    _print("mirror particle masks at start")
    lat = _lattice(cfg, parts, ri, bi, geo)
    st = _oracle(cfg, parts, ri, bi, geo)
    m = lat.masks()
    bad = [k for k in ("walls", "walls_all", "pin", "pout") if not np.array_equal(m[k], st.arrays[k])]
    assert bad == []
    # the spheres on the mirror planes reach into the ghost layers
    assert np.count_nonzero(m["walls"][-1, :, :] == 1) > 0 and np.count_nonzero(m["walls"][:, 0, :] == 2) > 0


@pytest.mark.parametrize("layout", ["rb", "fs"])
@pytest.mark.parametrize("nsteps", [1, 12])
def test_mirror_particle_steps_f64(nsteps, layout):
    cfg, parts, ri, bi, geo = _setup()
    _print("mirror particle f64 steps", nsteps, layout)
    lat = _lattice(cfg, parts, ri, bi, geo, layout=layout)
    st = _oracle(cfg, parts, ri, bi, geo)
    lat.step(nsteps)
    st.step(nsteps, particles=True)
    got = lat.get_particles()
    npar = len(parts["rad"])
    worst = {}
    for k, ok in (("center", "center"), ("up", "up"), ("rp", "rp"), ("force", "fp"), ("torque", "tp"),
                  ("fcap", "tension"), ("tcap", "tension_t"), ("frep", "frep"), ("shb", "shb")):
        ref = st.arrays[ok][:npar]
        d = np.max(np.abs(got[k] - ref))
        scale = max(np.max(np.abs(ref)), 1e-12)
        _print(k, "max abs diff", d, "scale", scale)
        worst[k] = d <= 1e-9 * scale + 1e-12
    m = lat.masks()
    bad = [k for k in ("walls", "walls_all", "pin", "pout") if not np.array_equal(m[k], st.arrays[k])]
    for k in bad:
        _print("mask", k, "mismatches", int(np.count_nonzero(m[k] != st.arrays[k])))
    assert bad == [], bad
    assert all(worst.values()), worst
    # the rules hold the plane particles on their planes
    assert abs(got["center"][0, 0] - (cfg["nx"] - 0.5)) < 1e-12 and abs(got["center"][1, 1] - 1.5) < 1e-12
    st.call("calmacros")
    mac = lat.calmacros()
    live = _live(ri.shape[:3])
    fluid = st.walls_all[live] == 0
    for k in MACROS:
        d = np.max(np.abs(mac[k][live][fluid] - st.arrays[k][live][fluid]))
        _print(k, "max abs diff", d)
        assert d <= 1e-11, k


def test_mirror_particles_refill_happens():
    """The pushed spheres uncover nodes, next to the mirror planes too."""
    cfg, parts, ri, bi, geo = _setup()
    st = _oracle(cfg, parts, ri, bi, geo)
    refilled = 0
    for _ in range(6):
        st.step(1, particles=True)
        refilled += int(np.count_nonzero((st.pout >= 1) & (st.walls_old >= 1)))
    assert refilled > 0


@pytest.mark.parametrize("layout", ["rb", "fs"])
def test_mirror_particles_f32_close_to_f64(layout):
    cfg, parts, ri, bi, geo = _setup()
    _print("mirror particle f32 vs f64", layout)
    res = {}
    for prec in ("f64", "f32"):
        lat = _lattice(cfg, parts, ri, bi, geo, prec, layout="rb" if prec == "f64" else layout)
        lat.step(30)
        res[prec] = (lat.calmacros(), lat.get_particles(), lat.masks())
    assert np.array_equal(res["f32"][2]["walls_all"], res["f64"][2]["walls_all"])
    live = _live(ri.shape[:3])
    fluid = res["f64"][2]["walls_all"][live] == 0
    a, b = res["f32"][0]["phi"][live][fluid], res["f64"][0]["phi"][live][fluid]
    l2 = np.linalg.norm(a - b) / np.linalg.norm(b)
    dc = np.max(np.abs(res["f32"][1]["center"] - res["f64"][1]["center"]))
    _print("phi l2", l2, "center diff", dc)
    assert l2 < 1e-5 and dc < 1e-5
