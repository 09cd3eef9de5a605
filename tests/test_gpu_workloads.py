"""GPU: the particle and ensemble workloads at their own sizes.

* Single sphere at a flat interface (96^3, walls at z = 1 and z = 96,
  periodic in x and y, red below z = 48, sphere R=10 at (48, 48, 48),
  density 1.5, contact angle 60 degrees, released at rest; tau=0.8,
  A=0.005, beta=0.7): FP64 against the oracle for 20 steps (masks bit for
  bit, fields to 1e-11, particle state to 1e-9 of its scale), then FP32
  against FP64 over 2000 steps.
* Three spheres of different densities (0.8, 1.4, 2.0) and contact angles
  (30, 90 and 150 degrees) pushed across the interface of a 48^3 box with z
  walls: FP64 against the oracle, masks, fields and every particle quantity.
* The batched ensemble (64 periodic Rayleigh-Plateau boxes drawn with seed
  7): the FP32 batch equals every box run alone in FP32 bit for bit, the FP64
  batch every box alone in FP64 (itself bit-exact against the oracle, which
  the smallest box checks here), and FP32 against FP64 per box after 200
  steps.
* The dense particle-laden interface (512^3, 200 spheres R=8 from seed 99,
  densities in [0.8, 2.0], contact angles in [30, 150] degrees): the masks of
  set_walls (S:326-395) against a host brute force over each particle's
  bounding box at the start and after 10 steps, and FP32 against FP64 at
  full size (fields at fluid nodes and the 200 trajectories) after 100
  steps.

The particle masks on periodic axes take minimum-image displacements (each
particle's bounding box in unwrapped coordinates, wrapped into the box),
where the reference's box has no periodic axis; the brute force here
follows that rule.
"""
import math

import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import oracle_from, rel_err
from tests.test_gpu_ext import _live, _oracle_particles

pytestmark = pytest.mark.gpu

MACROS = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")
PART_KEYS = (("center", "center"), ("up", "up"), ("rp", "rp"), ("force", "fp"), ("torque", "tp"),
             ("fcap", "tension"), ("tcap", "tension_t"), ("frep", "frep"), ("shb", "shb"))


def _print(*a):
    print(*a, flush=True)


def _particle_lattice(cfg, parts, precision, ri=None, bi=None):
    n = cfg["nx"]
    lat = symmetry_amd.Lattice(n, cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"], rho0=cfg["rho0"],
                               nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], bc=cfg["bc"],
                               npar=len(parts["rad"]))
    if ri is None:
        lat.init_interface(cfg["z0"], 6.0 * cfg["beta"] * 0.134)
    else:
        lat.upload(ri, bi)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"],
                      up=parts.get("up"), rp=parts.get("rp"))
    return lat


def _check_against_oracle(lat, st, npar, ri, cfg, tag):
    got = lat.get_particles()
    for k, ok in PART_KEYS:
        a, b = got[k], st.arrays[ok][:npar]
        scale = max(float(np.max(np.abs(b))), 1e-12)
        d = float(np.max(np.abs(a - b)))
        _print(tag, k, "max |d|", d, "scale", scale)
        assert d <= 1e-9 * scale + 1e-12, (tag, k)
    m = lat.masks()
    for k in ("walls", "walls_all", "pin", "pout"):
        assert np.array_equal(m[k], st.arrays[k]), (tag, k, int(np.count_nonzero(m[k] != st.arrays[k])))
    st.call("calmacros")
    mac = lat.calmacros()
    live = _live(ri.shape[:3], cfg["bc"])
    fluid = st.walls_all[live] == 0
    for k in MACROS:
        d = float(np.max(np.abs(mac[k][live][fluid] - st.arrays[k][live][fluid])))
        assert d <= 1e-11, (tag, k, d)


# ------------------------------------------------------ single sphere, 96^3

def _config2():
    cfg = ic.particle_config2()
    parts = ic.sphere_particles([cfg["center"]], cfg["rad"], cfg["rho_p"], cfg["cita_deg"], cfg["sigma"])
    parts["up"] = np.zeros((1, 3))
    parts["rp"] = np.zeros((1, 3))
    n = cfg["nx"]
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    return cfg, parts, ri, bi


def test_config2_sphere_released_at_rest_vs_oracle():
    cfg, parts, ri, bi = _config2()
    _print("96^3 sphere R=10 at rest: f64 vs oracle, 20 steps")
    lat = _particle_lattice(cfg, parts, "f64", ri, bi)
    st = _oracle_particles(cfg, parts, ri, bi)
    lat.step(20)
    st.step(20, particles=True)
    assert np.max(np.abs(st.center[0] - cfg["center"])) > 1e-6  # the sphere has moved
    _check_against_oracle(lat, st, 1, ri, cfg, "config2")


def test_config2_f32_vs_f64_first_step():
    """The reference's wetting correction (S:536-587) guards theta -> 0 but not
    theta -> pi.  At the fluid node straight above the sphere's north pole the
    colour gradient points into the sphere, antiparallel to the wall normal,
    so sin(theta) is a rounding residue and the corrected normal takes the
    size 1/sin(theta) times the particle centre's off-axis rounding noise
    (about 5e6 in FP64 after one step).  Recolouring then writes colour
    populations of order one into a pure-blue node, and the curvature force
    of its neighbours reads the same normal.  The FP64 path reproduces this
    bit for bit; FP32 gets another rounding residue there.  Outside two
    links of the pole node, as far as one step carries it, the two
    precisions agree to FP32 rounding: measured phi 4.6e-7, rho 6e-8,
    u 8e-10."""
    cfg, parts, _, _ = _config2()
    _print("96^3 sphere: f32 vs f64, one step")
    res = {}
    for prec in ("f64", "f32"):
        lat = _particle_lattice(cfg, parts, prec)
        lat.step(1)
        res[prec] = (lat.calmacros(), lat.masks())
        lat.close()
    (m64, k64), (m32, k32) = res["f64"], res["f32"]
    assert np.array_equal(k64["walls_all"], k32["walls_all"])
    fluid = k64["walls_all"] == 0
    pole = np.rint(np.array(cfg["center"]) - 1.0 + np.array([0.0, 0.0, cfg["rad"] + 1.0])).astype(int)
    cheb = np.max(np.abs(np.indices(fluid.shape) - pole[:, None, None, None]), axis=0)
    idx = np.argwhere(fluid & (np.abs(m64["phi"]) > 1.0 + 1e-9))
    _print("FP64 |phi| > 1 at", len(idx), "nodes, at most", int(cheb[tuple(idx.T)].max()) if len(idx) else None,
           "links from the pole node", tuple(pole))
    assert len(idx) > 0 and cheb[tuple(idx.T)].max() <= 1
    keep = fluid & (cheb > 2)
    for k, tol in (("phi", 2e-6), ("rho", 2e-7), ("ux", 1e-8), ("uy", 1e-8), ("uz", 1e-8)):
        d = float(np.max(np.abs(m32[k][keep] - m64[k][keep])))
        _print("  away from the pole", k, "max |d|", d)
        assert d < tol, (k, d)


def test_config2_f32_vs_f64_long_run():
    """2000 steps from the reference's start.  The pole disturbance described
    in the one-step test is seeded by rounding noise and spreads, so the two
    precisions share the motion but not its last digits: measured centre
    difference 5 % of the displacement, velocity 8 %."""
    cfg, parts, _, _ = _config2()
    _print("96^3 sphere: f32 vs f64 over 2000 steps")
    res = {}
    for prec in ("f64", "f32"):
        lat = _particle_lattice(cfg, parts, prec)
        lat.step(2000)
        res[prec] = (lat.calmacros(), lat.get_particles(), lat.masks())
        lat.close()
    m64, p64, k64 = res["f64"]
    m32, p32, k32 = res["f32"]
    fluid = (k64["walls_all"] == 0) & (k32["walls_all"] == 0)
    e_phi = rel_err(m32["phi"][fluid], m64["phi"][fluid])
    e_u = rel_err(np.stack([m32[k][fluid] for k in ("ux", "uy", "uz")]),
                  np.stack([m64[k][fluid] for k in ("ux", "uy", "uz")]))
    d64 = p64["center"][0] - np.array(cfg["center"])
    d32 = p32["center"][0] - np.array(cfg["center"])
    dc = float(np.linalg.norm(d32 - d64))
    moved = float(np.linalg.norm(d64))
    du = float(np.linalg.norm(p32["up"][0] - p64["up"][0]) / np.linalg.norm(p64["up"][0]))
    _print("2000 steps: phi", e_phi, "u", e_u, "centre |d|", dc, "of a displacement", moved, "velocity rel", du,
           "d64", d64, "d32", d32)
    for m in (m32, m64):
        assert all(np.isfinite(m[k][fluid]).all() for k in ("phi", "ux", "uy", "uz"))
    assert moved > 1.0
    assert float(np.dot(d32, d64)) > 0.95 * moved * float(np.linalg.norm(d32))
    assert dc < 0.1 * moved and du < 0.2


# ----------------------------------- several spheres, densities and angles

def test_mixed_densities_and_angles_vs_oracle():
    n = 48
    cfg = ic.particle_config2(n=n, rad=5.0)
    centers = [(14.3, 15.2, 24.4), (33.6, 16.1, 23.7), (24.2, 33.4, 24.9)]
    parts = ic.sphere_particles(centers, 5.0, [0.8, 1.4, 2.0], [30.0, 90.0, 150.0], cfg["sigma"])
    parts["up"] = np.array([[0.02, -0.01, 0.03], [-0.02, 0.01, -0.03], [0.0, 0.02, 0.02]])
    parts["rp"] = np.array([[0.0, 0.0, 0.01], [0.01, 0.0, 0.0], [0.0, -0.01, 0.0]])
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    for nsteps in (1, 10):
        _print("mixed particles f64 vs oracle", nsteps, "steps")
        lat = _particle_lattice(cfg, parts, "f64", ri, bi)
        st = _oracle_particles(cfg, parts, ri, bi)
        lat.step(nsteps)
        st.step(nsteps, particles=True)
        _check_against_oracle(lat, st, 3, ri, cfg, f"mixed {nsteps}")
        lat.close()


# ----------------------------------------------------- batched ensemble

KPROF = 6.0 * 0.7 * 0.134


def _ensemble_batch(boxes, precision, layout):
    bt = symmetry_amd.Batch([(b["nx"], b["ny"], b["nz"], b["nu_r"], b["nu_b"]) for b in boxes], precision=precision,
                            layout=layout, sigma=0.01, rho0=1.0, beta=0.7, xchunk=8)
    for k, b in enumerate(boxes):
        bt.init_cylinder(k, b["nx"] / 2 + 0.5, b["ny"] / 2 + 0.5, b["r0"], KPROF, b["amps"])
    return bt


def _single(b, precision, layout):
    lat = symmetry_amd.Lattice(b["nx"], b["ny"], b["nz"], precision=precision, sigma=0.01, rho0=1.0, nu_r=b["nu_r"],
                               nu_b=b["nu_b"], beta=0.7, layout=layout, xchunk=8)
    lat.init_cylinder(b["nx"] / 2 + 0.5, b["ny"] / 2 + 0.5, b["r0"], KPROF, b["amps"])
    return lat


@pytest.mark.parametrize("precision,layout", [("f32", "fs"), ("f64", "rb")])
def test_config4_batch_equals_every_box_alone(precision, layout):
    boxes = ic.ensemble_config4()
    assert len(boxes) == 64
    nsteps = 5
    _print("64-box ensemble", precision, layout, nsteps, "steps")
    bt = _ensemble_batch(boxes, precision, layout)
    if precision == "f32":
        assert bt.scatter_form()
    bt.step(nsteps)
    for k, b in enumerate(boxes):
        lat = _single(b, precision, layout)
        lat.step(nsteps)
        ra, ba = bt.download(k)
        rs, bs = lat.download()
        lat.close()
        assert np.array_equal(ra, rs) and np.array_equal(ba, bs), (k, b["nx"], b["nz"])
    if precision == "f64":
        # the smallest box against the oracle (single boxes are pinned there too)
        k = min(range(64), key=lambda j: boxes[j]["nx"] * boxes[j]["ny"] * boxes[j]["nz"])
        b = boxes[k]
        ri, bi, _ = ic.workload_state(b, exact=False)
        bt2 = _ensemble_batch(boxes, "f64", "rb")
        bt2.upload(k, ri, bi)
        bt2.step(3)
        st = oracle_from(b, ri, bi)
        st.step(3)
        r2, b2 = bt2.download(k)
        assert np.array_equal(r2, st.ri) and np.array_equal(b2, st.bi), k


def test_config4_f32_vs_f64_every_box():
    """FP32 factorised batch against FP64 per box after 200 steps.  phi
    agrees to 2e-7; the velocity, still small at step 200, carries the
    populations' FP32 rounding relative to its own size: measured worst
    3.5e-5 relative (L2) in the boxes with tau near 0.6, 1e-6 to 1e-5 in
    the rest."""
    boxes = ic.ensemble_config4()
    nsteps = 200
    _print("64-box ensemble f32 vs f64,", nsteps, "steps")
    b32 = _ensemble_batch(boxes, "f32", "fs")
    b64 = _ensemble_batch(boxes, "f64", "rb")
    b32.step(nsteps)
    b64.step(nsteps)
    worst = (0.0, 0.0)
    for k in range(len(boxes)):
        m32, m64 = b32.calmacros(k), b64.calmacros(k)
        e_phi = rel_err(m32["phi"], m64["phi"])[0]
        e_u = rel_err(np.stack([m32[c] for c in ("ux", "uy", "uz")]), np.stack([m64[c] for c in ("ux", "uy", "uz")]))[0]
        _print("box", k, boxes[k]["nx"], boxes[k]["nz"], "tau %.3f" % boxes[k]["tau"], "phi %.2e u %.2e" % (e_phi, e_u))
        worst = (max(worst[0], e_phi), max(worst[1], e_u))
    _print("worst over the boxes: phi %.2e u %.2e" % worst)
    assert worst[0] < 1e-5 and worst[1] < 1e-4


# ------------------------------------------- dense particle-laden interface

def _brute_masks(n, centers, rad, zwalls=True):
    """set_walls (S:326-395) by brute force over each particle's bounding box
    in unwrapped coordinates (minimum image on the periodic x and y axes),
    highest particle index winning where spheres share a node; z walls with
    ghost layers at z = 1 and z = n."""
    N = n ** 3
    walls = np.zeros(N, dtype=np.int32)
    pin = np.zeros(N, dtype=np.int32)
    pout = np.zeros(N, dtype=np.int32)
    C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                  [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1], [1, 0, 1],
                  [-1, 0, -1], [1, 0, -1], [-1, 0, 1]])

    def box(c, r, m):
        lo = [int(math.floor(c[a] - 1.0 - r - m)) for a in range(3)]
        hi = [int(math.ceil(c[a] - 1.0 + r + m)) for a in range(3)]
        g = np.meshgrid(*[np.arange(lo[a], hi[a] + 1) for a in range(3)], indexing="ij")
        return [x.ravel() for x in g]

    def index(X, Y, Z):
        return ((X % n) * n + (Y % n)) * n + (Z % n)

    for p, c in enumerate(centers):
        X, Y, Z = box(c, rad, 3.0)
        keep = (Z >= 0) & (Z < n)
        X, Y, Z = X[keep], Y[keep], Z[keep]
        dx, dy, dz = (X + 1.0) - c[0], (Y + 1.0) - c[1], (Z + 1.0) - c[2]
        solid = dx * dx + dy * dy + dz * dz <= rad * rad
        idx = index(X[solid], Y[solid], Z[solid])
        walls[idx] = np.maximum(walls[idx], p + 1)
    solid_all = walls > 0
    for p, c in enumerate(centers):
        X, Y, Z = box(c, rad, 1.0)
        keep = (Z >= 1) & (Z <= n - 2)
        X, Y, Z = X[keep], Y[keep], Z[keep]
        nidx = index(X, Y, Z)
        live = ~solid_all[nidx]
        X, Y, Z, nidx = X[live], Y[live], Z[live], nidx[live]
        for i in range(1, 19):
            m = index(X + C[i, 0], Y + C[i, 1], Z + C[i, 2])
            hit = walls[m] == p + 1
            pout[nidx[hit]] = np.maximum(pout[nidx[hit]], p + 1)
            pin[m[hit]] = np.maximum(pin[m[hit]], p + 1)
    shape = (n, n, n)
    return {"walls": walls.reshape(shape), "walls_all": np.where(solid_all, 10, 0).astype(np.int32).reshape(shape),
            "pin": pin.reshape(shape), "pout": pout.reshape(shape)}


def _config5_lattice(precision):
    c5 = ic.particle_config5()
    parts = ic.sphere_particles(c5["centers"], c5["rad"], c5["rho_p"], c5["cita_deg"], c5["sigma"])
    return c5, _particle_lattice(c5, parts, precision)


def test_config5_masks_full_size():
    c5, lat = _config5_lattice("f32")
    n = c5["nx"]
    for nsteps in (0, 10):
        if nsteps:
            lat.step(nsteps)
        centers = lat.get_particles()["center"]
        _print("512^3 masks after", nsteps, "steps")
        ref = _brute_masks(n, centers, c5["rad"])
        got = lat.masks()
        for k in ("walls", "walls_all", "pin", "pout"):
            bad = int(np.count_nonzero(got[k] != ref[k]))
            _print(k, "nonzero", int(np.count_nonzero(ref[k])), "mismatches", bad)
            assert bad == 0, k
        wrapped = [c for c in centers if min(c[0], c[1]) < c5["rad"] + 1 or max(c[0], c[1]) > n - c5["rad"]]
        assert wrapped  # spheres across the periodic faces are part of the check
    lat.close()


def test_config5_f32_vs_f64_full_size():
    nsteps = 100
    res = {}
    for prec in ("f64", "f32"):
        _print("512^3, 200 spheres,", prec, nsteps, "steps")
        c5, lat = _config5_lattice(prec)
        lat.step(nsteps)
        m = lat.calmacros()
        res[prec] = ({k: m[k] for k in ("phi", "ux", "uy", "uz")}, lat.get_particles(), lat.masks()["walls_all"])
        del m
        lat.close()
    (m64, p64, w64), (m32, p32, w32) = res["f64"], res["f32"]
    same = int(np.count_nonzero(w64 != w32))
    fluid = (w64 == 0) & (w32 == 0)
    e_phi = rel_err(m32["phi"][fluid], m64["phi"][fluid])
    e_u = rel_err(np.stack([m32[k][fluid] for k in ("ux", "uy", "uz")]),
                  np.stack([m64[k][fluid] for k in ("ux", "uy", "uz")]))
    moved = np.max(np.abs(p64["center"] - np.asarray(ic.particle_config5()["centers"])), axis=1)
    dc = np.max(np.abs(p32["center"] - p64["center"]), axis=1)
    du = np.max(np.abs(p32["up"] - p64["up"]), axis=1) / np.maximum(np.max(np.abs(p64["up"]), axis=1), 1e-30)
    _print("masks differing", same, "phi", e_phi, "u", e_u, "centre |d| max", float(dc.max()),
           "displacement max", float(moved.max()), "velocity rel max", float(du.max()))
    assert same == 0
    assert e_phi[0] < 1e-5 and e_u[0] < 1e-4
    assert float(dc.max()) < 1e-6 and float(np.median(du)) < 1e-3
