"""GPU parity with wall faces and moving particles, against the FP64 oracle.

* Wall faces only: the FP64 library reproduces the oracle bit for bit
  (half-way bounce-back is a copy, S:1544-1592), in the red/blue layout and
  in the factorised layout (whose solid nodes keep their slots in red/blue
  side arrays).
* Moving particles: masks (walls, walls_all, pin, pout; set_walls S:326-395)
  must match exactly; fields and particle trajectories match within a
  tolerance, because the FP64 force/torque sums are block reductions on the
  GPU and serial sums in the oracle (S:1178-1261, S:1998-2022): their
  rounding differs in the last bits, and that difference is what the
  tolerances bound.
"""
import math

import numpy as np
import pytest

import oracle
import symmetry_amd
from symmetry_amd import ic
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

MACROS = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")


def _print(*a):
    print(*a, flush=True)


def _live(shape, bc):
    """Slices of the nodes that are not wall or mirror ghost layers."""
    sl = []
    for a, n in enumerate(shape):
        lo = 1 if bc[2 * a] != symmetry_amd.BC_PERIODIC else 0
        hi = n - 1 if bc[2 * a + 1] != symmetry_amd.BC_PERIODIC else n
        sl.append(slice(lo, hi))
    return tuple(sl)


def _lattice(cfg, precision="f64", **kw):
    return symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                                rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"],
                                bc=cfg["bc"], **kw)


def _oracle(cfg, ri, bi, npar=0, fg=0.0):
    st = oracle.OracleState(cfg["nx"], cfg["ny"], cfg["nz"], bc=cfg["bc"], npar=npar, sigma=cfg["sigma"],
                            rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], fg=fg)
    st.ri[...] = ri
    st.bi[...] = bi
    return st


def _wall_cfg(nx, ny, nz, bc):
    return dict(nx=nx, ny=ny, nz=nz, bc=bc, sigma=0.01, rho0=1.0, nu_r=0.1, nu_b=1.0 / 6.0, beta=0.7)


WALL_CASES = {
    # walls at z = 1 and z = nz, periodic x / y (the particle workload's box)
    "z": (_wall_cfg(32, 24, 40, (0, 0, 0, 0, 1, 1)), (16.0, 12.0, 7.0, 8.0)),
    # a closed box
    "xyz": (_wall_cfg(20, 24, 28, (1, 1, 1, 1, 1, 1)), (6.0, 7.0, 8.0, 6.5)),
    # walls in x only
    "x": (_wall_cfg(24, 16, 32, (1, 1, 0, 0, 0, 0)), (4.0, 8.5, 16.0, 6.0)),
    # the reference's own faces: walls at x=1 and z=1, mirrors at x=xl, y=1,
    # y=yl and z=zl (S:806-807, S:1443-1592); the drop sits on the mirrors
    "ref": (_wall_cfg(20, 18, 24, (1, 2, 2, 2, 1, 2)), (18.0, 2.0, 12.0, 6.0)),
    # mirrors all round, drop in a corner
    "mirror": (_wall_cfg(16, 20, 22, (2, 2, 2, 2, 2, 2)), (3.0, 17.0, 4.0, 5.5)),
    # mirrors in x, walls in z, periodic y
    "mixed": (_wall_cfg(18, 16, 26, (2, 2, 0, 0, 1, 1)), (16.5, 8.0, 3.0, 6.0)),
}


@pytest.mark.parametrize("case", sorted(WALL_CASES))
def test_wall_roundtrip(case):
    cfg, drop = WALL_CASES[case]
    _print("wall roundtrip", case)
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=-1e3, rho0=1.0, drop=drop,
                                        sigma=cfg["sigma"])
    lat = _lattice(cfg, layout="fs")
    lat.upload(ri, bi)
    r2, b2 = lat.download()
    live = _live(ri.shape[:3], cfg["bc"])
    assert np.array_equal(r2[live], ri[live]) and np.array_equal(b2[live], bi[live])


@pytest.mark.parametrize("layout", ["rb", "fs"])
@pytest.mark.parametrize("nsteps", [1, 15])
@pytest.mark.parametrize("case", sorted(WALL_CASES))
def test_walls_f64_bit_exact(case, nsteps, layout):
    """A red drop touching the wall(s), relaxing for nsteps steps."""
    cfg, drop = WALL_CASES[case]
    _print("walls f64", case, nsteps, layout)
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=-1e3, rho0=1.0, drop=drop,
                                        sigma=cfg["sigma"])
    lat = _lattice(cfg, layout=layout)
    lat.upload(ri, bi)
    lat.step(nsteps)
    r2, b2 = lat.download()
    st = _oracle(cfg, ri, bi)
    st.step(nsteps)
    live = _live(ri.shape[:3], cfg["bc"])
    dr = np.max(np.abs(r2[live] - st.ri[live]))
    db = np.max(np.abs(b2[live] - st.bi[live]))
    _print("max |dr|, |db|:", dr, db)
    assert dr == 0.0 and db == 0.0
    st.call("calmacros")
    mac = lat.calmacros()
    for k in MACROS:
        assert np.array_equal(mac[k][live], st.arrays[k][live]), k
    # ghost layers of phi hold the adjacent layer (S:806-807)
    assert np.array_equal(mac["phi"], st.phi)


@pytest.mark.parametrize("layout", ["rb", "fs"])
@pytest.mark.parametrize("case", ["ref", "mirror", "z"])
def test_wall_derivatives_match_oracle(case, layout):
    """Gradient, normal and flag of the derivatives stage (S:452-597) at every
    node, ghost layers included: zero at wall ghosts, the reflected copy of
    the image at mirror ghosts (S:589-594)."""
    cfg, drop = WALL_CASES[case]
    _print("wall derivatives", case, layout)
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=-1e3, rho0=1.0, drop=drop,
                                        sigma=cfg["sigma"])
    lat = _lattice(cfg, layout=layout)
    lat.upload(ri, bi)
    lat.step(5)
    st = _oracle(cfg, ri, bi)
    st.step(5)
    st.call("calmacros")
    st.call("derivatives")
    der = lat.derivatives()
    for k in ("gx", "gy", "gz", "nxv", "nyv", "nzv"):
        assert np.array_equal(der[k], st.arrays[k]), k
    assert np.array_equal(der["flag"], st.flag)


@pytest.mark.parametrize("layout", ["rb", "fs"])
def test_walls_f32_close_to_f64(layout):
    """A perturbed flat interface between z walls relaxing as a capillary
    wave, 200 steps: FP32 phi within 1e-5 (L2-relative) of the FP64 path and
    velocity within 3e-5.  The wave's velocities are ~1e-3, so the FP32
    momentum sum's absolute rounding (~1e-8 per node) is a larger share of
    them than in the Rayleigh-Plateau workload.  The geometry keeps phi monotone in z and every node within
    12 nodes of the interface.  Where grad phi vanishes in a mixed region (a
    saddle of phi between a drop and its periodic image) the colour-gradient
    normal n = grad phi / |grad phi| is ill-conditioned and FP32 and FP64
    legitimately part; where a colour density falls below the evaporation
    limit 1e-8 (S:716-725) the recolouring switches discontinuously."""
    cfg = _wall_cfg(32, 16, 24, (0, 0, 0, 0, 1, 1))
    _print("walls f32 vs f64", layout)
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=12.5, amp=2.0, rho0=1.0)
    out = {}
    for prec in ("f64", "f32"):
        lat = _lattice(cfg, prec, layout="rb" if prec == "f64" else layout)
        lat.upload(ri, bi)
        lat.step(200)
        out[prec] = lat.calmacros()
    live = _live(ri.shape[:3], cfg["bc"])
    l2, linf = rel_err(out["f32"]["phi"][live], out["f64"]["phi"][live])
    u = np.stack([out["f32"][k][live] for k in ("ux", "uy", "uz")])
    u64 = np.stack([out["f64"][k][live] for k in ("ux", "uy", "uz")])
    ul2, _ = rel_err(u, u64)
    _print("phi rel err", l2, linf, "u rel err", ul2)
    assert l2 < 1e-5 and ul2 < 3e-5


# ------------------------------------------------------------------ particles

def _particle_setup(n=40, rad=6.0, up=(0.3, 0.2, -0.05), rp=(0.0, 0.0, 0.02), offset=(0.3, -0.2, 0.4)):
    cfg = ic.particle_config2(n=n, rad=rad)
    c = np.array(cfg["center"]) + np.array(offset)
    parts = ic.sphere_particles([c], rad, cfg["rho_p"], cfg["cita_deg"], cfg["sigma"])
    parts["up"] = np.array([up], dtype=np.float64)
    parts["rp"] = np.array([rp], dtype=np.float64)
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    return cfg, parts, ri, bi


def _oracle_particles(cfg, parts, ri, bi):
    """The reference's start-up (S:209-214): masks, state, derivatives."""
    npar = len(parts["rad"])
    st = _oracle(cfg, ri, bi, npar=npar, fg=parts["fg"])
    st.center[:npar] = parts["center"]
    st.center_old[:npar] = parts["center"]
    st.center0[:npar] = parts["center"]
    for k in ("rad", "mass", "inertia", "cita"):
        st.arrays[k][:npar] = parts[k]
    st.up[:npar] = parts["up"]
    st.rp[:npar] = parts["rp"]
    dims = (cfg["nx"], cfg["ny"], cfg["nz"])
    for a in range(3):
        for side in range(2):
            on = cfg["bc"][2 * a + side] == symmetry_amd.BC_WALL
            st.s.rep_on[2 * a + side] = int(on)
            st.s.rep_plane[2 * a + side] = 1.5 if side == 0 else dims[a] - 0.5
    st.call("set_walls")
    st.ffi[...] = st.ri + st.bi  # initialization's ffi = fi (S:894)
    st.call("calmacros")
    st.call("derivatives")
    return st


def _lattice_particles(cfg, parts, ri, bi, precision="f64", layout="rb"):
    lat = _lattice(cfg, precision, npar=len(parts["rad"]), layout=layout)
    lat.upload(ri, bi)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"],
                      up=parts["up"], rp=parts["rp"])
    return lat


def _masks_equal(lat, st):
    m = lat.masks()
    bad = [k for k in ("walls", "walls_all", "pin", "pout") if not np.array_equal(m[k], st.arrays[k])]
    return bad, m


def test_particle_masks_at_start():
    cfg, parts, ri, bi = _particle_setup()
    _print("particle masks at start")
    lat = _lattice_particles(cfg, parts, ri, bi)
    st = _oracle_particles(cfg, parts, ri, bi)
    bad, m = _masks_equal(lat, st)
    assert bad == []
    assert np.count_nonzero(m["pout"]) > 0 and np.count_nonzero(m["pin"]) > 0


@pytest.mark.parametrize("layout", ["rb", "fs"])
@pytest.mark.parametrize("nsteps", [1, 12])
def test_particle_steps_f64(nsteps, layout):
    """Moving sphere (pushed so that nodes are covered and uncovered) at a
    flat interface between z walls: masks exact, fields and trajectory to
    rounding of the force sums."""
    cfg, parts, ri, bi = _particle_setup()
    _print("particle f64 steps", nsteps, layout)
    lat = _lattice_particles(cfg, parts, ri, bi, layout=layout)
    st = _oracle_particles(cfg, parts, ri, bi)
    lat.step(nsteps)
    st.step(nsteps, particles=True)
    got = lat.get_particles()
    worst = {}
    for k, ok in (("center", "center"), ("up", "up"), ("rp", "rp"), ("force", "fp"), ("torque", "tp"),
                  ("fcap", "tension"), ("tcap", "tension_t"), ("frep", "frep"), ("shb", "shb")):
        d = np.max(np.abs(got[k] - st.arrays[ok][:1]))
        scale = max(np.max(np.abs(st.arrays[ok][:1])), 1e-12)
        _print(k, got[k][0], st.arrays[ok][0], "rel", d / scale)
        worst[k] = d <= 1e-9 * scale + 1e-12  # sums of O(0.1) link terms
    bad, m = _masks_equal(lat, st)
    for k in bad:
        _print("mask", k, "mismatches", int(np.count_nonzero(m[k] != st.arrays[k])))
    assert bad == [], bad
    assert all(worst.values()), worst
    st.call("calmacros")
    mac = lat.calmacros()
    live = _live(ri.shape[:3], cfg["bc"])
    fluid = st.walls_all[live] == 0
    for k in MACROS:
        a, b = mac[k][live][fluid], st.arrays[k][live][fluid]
        d = np.max(np.abs(a - b))
        _print(k, "max abs diff", d)
        assert d <= 1e-11, k
    r2, b2 = lat.download()
    sel = (st.walls_all == 0)[..., None] & np.ones(19, bool)
    sel[:, :, 0] = sel[:, :, -1] = False  # wall ghost layers
    assert np.max(np.abs(r2[sel] - st.ri[sel])) <= 1e-11
    assert np.max(np.abs(b2[sel] - st.bi[sel])) <= 1e-11


def test_particle_refill_happens():
    """The pushed sphere uncovers nodes within the first steps, so the refill
    path (S:1351-1418) is part of the compared run."""
    cfg, parts, ri, bi = _particle_setup()
    st = _oracle_particles(cfg, parts, ri, bi)
    refilled = 0
    for _ in range(4):
        st.step(1, particles=True)
        refilled += int(np.count_nonzero((st.pout >= 1) & (st.walls_old >= 1)))
    assert refilled > 0


@pytest.mark.parametrize("layout", ["rb", "fs"])
def test_particle_f32_close_to_f64(layout):
    cfg, parts, ri, bi = _particle_setup()
    _print("particle f32 vs f64", layout)
    res = {}
    for prec in ("f64", "f32"):
        lat = _lattice_particles(cfg, parts, ri, bi, prec, layout="rb" if prec == "f64" else layout)
        lat.step(30)
        res[prec] = (lat.calmacros(), lat.get_particles(), lat.masks())
    assert np.array_equal(res["f32"][2]["walls_all"], res["f64"][2]["walls_all"])
    fluid = res["f64"][2]["walls_all"] == 0
    l2, linf = rel_err(res["f32"][0]["phi"][fluid], res["f64"][0]["phi"][fluid])
    _print("phi", l2, linf)
    assert l2 < 1e-5
    dc = np.max(np.abs(res["f32"][1]["center"] - res["f64"][1]["center"]))
    _print("center diff", dc)
    assert dc < 1e-5


@pytest.mark.parametrize("layout", ["rb", "fs"])
def test_particle_trajectory_vs_oracle_and_f32(layout):
    """A sphere released at rest at the interface (offset from the lattice so
    no node sits exactly on its surface): FP64 centre and velocity against the
    oracle after 50 steps, FP32 against FP64 after 300 steps."""
    cfg, parts, ri, bi = _particle_setup(up=(0.0, 0.0, 0.0), rp=(0.0, 0.0, 0.0))
    _print("particle trajectory", layout)
    st = _oracle_particles(cfg, parts, ri, bi)
    st.step(50, particles=True)
    traj = {}
    for prec in ("f64", "f32"):
        lat = _lattice_particles(cfg, parts, ri, bi, prec, layout=layout)
        lat.step(50)
        traj[prec, 50] = lat.get_particles()
        lat.step(250)
        traj[prec, 300] = lat.get_particles()
        lat.close()
    g = traj["f64", 50]
    dc = np.max(np.abs(g["center"][0] - st.center[0]))
    du = np.max(np.abs(g["up"][0] - st.up[0])) / np.max(np.abs(st.up[0]))
    _print("f64 vs oracle at 50: centre", dc, "velocity rel", du)
    assert dc < 1e-9 and du < 1e-6
    a, b = traj["f32", 300], traj["f64", 300]
    dc = np.max(np.abs(a["center"][0] - b["center"][0]))
    du = np.max(np.abs(a["up"][0] - b["up"][0])) / np.max(np.abs(b["up"][0]))
    _print("f32 vs f64 at 300: centre", dc, "velocity rel", du)
    assert dc < 1e-5 and du < 1e-3


@pytest.mark.parametrize("case", ["z", "ref"])
def test_device_interface_ic_matches_host(case):
    """cgl_init_interface equals the host flat-interface state (uploaded) at
    every live node to rounding of tanh, ghost layers included in the pull
    layout's bounce-back and mirror slots."""
    cfg, _ = WALL_CASES[case]
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    kprof = 6.0 * cfg["beta"] * 0.134
    ri, bi, _ = ic.flat_interface_state(nx, ny, nz, z0=nz / 2.0, rho0=cfg["rho0"], beta=cfg["beta"])
    lat = _lattice(cfg)
    lat.init_interface(nz / 2.0, kprof)
    r2, b2 = lat.download()
    live = _live(ri.shape[:3], cfg["bc"])
    assert np.max(np.abs(r2[live] - ri[live])) < 1e-14
    assert np.max(np.abs(b2[live] - bi[live])) < 1e-14
