"""The per-stage entry points of a step with particles (cgl_capillary,
cgl_momentum, cgl_repulsive, cgl_move, cgl_set_walls, cgl_refill,
cgl_collide_stream) against the oracle's matching stage, one stage at a
time, and a staged step against a cgl_step step bit for bit.

Two cases: the pushed sphere of the particle tests (40^3, z walls, flat
interface, R = 6, velocity (0.3, 0.2, -0.05) per step, spinning), which
covers and uncovers nodes every step, so refilling has nodes to fill; and two
moving spheres of different densities and contact angles across the
interface, one within the repulsion range of the z = 1 wall and the pair
within the range of each other, so the repulsion stage has wall and pair
forces to compute."""
import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.test_gpu_ext import WALL_CASES, _lattice, _lattice_particles, _live, _oracle_particles, _particle_setup

pytestmark = pytest.mark.gpu

PART = {"center": "center", "up": "up", "rp": "rp", "force": "fp", "torque": "tp", "fcap": "tension",
        "tcap": "tension_t", "frep": "frep", "shb": "shb"}
MACROS = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")


def _print(*a):
    print(*a, flush=True)


def _particles_match(lat, st, keys, tag):
    got = lat.get_particles()
    for k in keys:
        a = got[k]
        b = st.arrays[PART[k]][:len(a)]
        scale = max(float(np.max(np.abs(b))), 1e-12)
        d = float(np.max(np.abs(a - b)))
        _print(tag, k, "max |d|", d, "scale", scale)
        assert d <= 1e-9 * scale + 1e-12, (tag, k, d)  # sums of O(0.1) link terms in another order


def _masks_match(lat, st, tag):
    m = lat.masks()
    for k in ("walls", "walls_all", "pin", "pout"):
        assert np.array_equal(m[k], st.arrays[k]), (tag, k, int(np.count_nonzero(m[k] != st.arrays[k])))


def _macros_match(mac, st, cfg, tag):
    live = _live(st.rho.shape, cfg["bc"])
    fluid = st.walls_all[live] == 0
    for k in MACROS:
        d = float(np.max(np.abs(mac[k][live][fluid] - st.arrays[k][live][fluid])))
        assert d <= 1e-11, (tag, k, d)


def _two_spheres_near_a_wall():
    n = 40
    cfg = ic.particle_config2(n=n, rad=5.0)
    cfg["z0"] = 11.0
    parts = ic.sphere_particles([(14.3, 20.2, 9.4), (25.6, 20.9, 10.1)], 5.0, [1.2, 1.8], [50.0, 120.0],
                                cfg["sigma"])
    parts["up"] = np.array([[-0.2, 0.1, -0.05], [0.25, -0.1, 0.1]])
    parts["rp"] = np.array([[0.0, 0.01, 0.0], [0.0, 0.0, -0.02]])
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=cfg["z0"], rho0=cfg["rho0"], beta=cfg["beta"])
    return cfg, parts, ri, bi


@pytest.mark.parametrize("case", ["pushed sphere", "two spheres near a wall"])
def test_staged_steps_match_the_oracle_stage_by_stage(case):
    cfg, parts, ri, bi = _particle_setup() if case == "pushed sphere" else _two_spheres_near_a_wall()
    lat = _lattice_particles(cfg, parts, ri, bi)
    ref = _lattice_particles(cfg, parts, ri, bi)
    st = _oracle_particles(cfg, parts, ri, bi)
    lat.step(2)
    ref.step(2)
    st.step(2, particles=True)
    refilled_total = 0
    for k in range(3):
        tag = f"staged step {k + 1}"
        _print(tag)
        mac = lat.calmacros()
        st.call("calmacros")
        _macros_match(mac, st, cfg, tag + " calmacros")
        lat.stage("capillary")
        st.call("lagrange_capillary")
        _particles_match(lat, st, ("fcap", "tcap", "shb"), tag + " capillary")
        lat.stage("momentum")
        st.call("momentum")
        _particles_match(lat, st, ("force", "torque"), tag + " momentum")
        lat.stage("repulsive")
        st.call("repulsive")
        _particles_match(lat, st, ("frep",), tag + " repulsive")
        if case != "pushed sphere" and k == 0:  # wall and pair repulsion on both spheres
            assert np.all(np.abs(st.frep[:2]).max(axis=1) > 1.0)
        lat.stage("move")
        st.call("movement")
        _particles_match(lat, st, ("center", "up", "rp", "force", "torque", "fcap", "tcap", "frep"), tag + " move")
        lat.stage("set_walls")
        st.call("set_walls")
        _masks_match(lat, st, tag + " set_walls")
        lat.stage("refill")
        st.call("refilling")
        live = np.zeros(st.rho.shape, bool)
        live[_live(st.rho.shape, cfg["bc"])] = True
        refilled = live & (st.pout >= 1) & (st.walls_old >= 1)
        refilled_total += int(np.count_nonzero(refilled))
        band = lat.particle_band()
        for f in ("rho", "rho_r", "rho_b", "ux", "uy", "uz"):
            if refilled.any():
                b = st.arrays[f][refilled]
                d = float(np.max(np.abs(band[f][refilled] - b)))
                assert d <= 1e-9 * max(float(np.max(np.abs(b))), 1e-12) + 1e-15, (tag, "refill", f, d)
        _print(tag, "refilled nodes", int(np.count_nonzero(refilled)))
        lat.stage("collide_stream")
        for fn in ("derivatives", "collision", "recolor", "boundary"):
            st.call(fn)
        st.streaming("ri")
        st.streaming("bi")
        st.call("curve")
    assert refilled_total > 0
    ref.step(3)
    st.call("calmacros")
    _macros_match(lat.calmacros(), st, cfg, "after the staged steps")
    ra, ba = lat.download()
    rb, bb = ref.download()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
    pa, pb = lat.get_particles(), ref.get_particles()
    for k in pa:
        assert np.array_equal(pa[k], pb[k]), k


def test_staged_step_without_particles_equals_a_step():
    cfg, drop = WALL_CASES["z"]
    ri, bi, _ = ic.flat_interface_state(cfg["nx"], cfg["ny"], cfg["nz"], z0=cfg["nz"] / 2.0, rho0=cfg["rho0"],
                                        beta=cfg["beta"], drop=drop)
    a, b = _lattice(cfg), _lattice(cfg)
    a.upload(ri, bi)
    b.upload(ri, bi)
    for _ in range(3):
        a.calmacros()
        a.stage("collide_stream")
    b.step(3)
    ra, ba = a.download()
    rb, bb = b.download()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
    with pytest.raises(symmetry_amd.CglError) as e:
        a.calmacros()
        a.stage("capillary")
    assert e.value.code == symmetry_amd.ERR_ARG


def test_stage_order_is_enforced():
    cfg, parts, ri, bi = _particle_setup()
    lat = _lattice_particles(cfg, parts, ri, bi)
    with pytest.raises(symmetry_amd.CglError) as e:
        lat.stage("capillary")  # no calmacros yet
    assert e.value.code == symmetry_amd.ERR_ARG
    lat.calmacros()
    lat.stage("capillary")
    for bad in (lambda: lat.stage("move"), lambda: lat.step(1), lambda: lat.calmacros(), lambda: lat.download(),
                lambda: lat.diagnostics(), lambda: lat.stage("collide_stream")):
        with pytest.raises(symmetry_amd.CglError) as e:
            bad()
        assert e.value.code == symmetry_amd.ERR_ARG
    lat.get_particles()  # read-backs of the particle state stay open
    lat.masks()
    for s in ("momentum", "repulsive", "move", "set_walls", "refill", "collide_stream"):
        lat.stage(s)
    lat.step(1)
    lat.calmacros()
    lat.stage("capillary")
    lat.upload(ri, bi)  # installing a state abandons the staged step
    lat.step(1)
