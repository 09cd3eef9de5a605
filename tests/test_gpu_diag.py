"""Device FP64 reductions (cgl_diagnostics) against host sums of the
downloaded macroscopic fields, over live fluid nodes only."""
import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import small_rp

pytestmark = pytest.mark.gpu


def _host_sums(mac, live):
    rho = mac["rho"][live]
    ux, uy, uz = mac["ux"][live], mac["uy"][live], mac["uz"][live]
    return {"mass": rho.sum(), "mass_r": mac["rho_r"][live].sum(), "mass_b": mac["rho_b"][live].sum(),
            "jx": (rho * ux).sum(), "jy": (rho * uy).sum(), "jz": (rho * uz).sum(),
            "ekin": (0.5 * rho * (ux * ux + uy * uy + uz * uz)).sum()}


def _check(dev, host, scale):
    for k, v in host.items():
        assert abs(dev[k] - v) <= 1e-12 * scale, (k, dev[k], v)


@pytest.mark.parametrize("precision,layout", [("f64", "rb"), ("f32", "fs")])
def test_diagnostics_periodic(precision, layout):
    cfg, ri, bi = small_rp()
    lat = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                               nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], layout=layout)
    lat.upload(ri, bi)
    d0 = lat.diagnostics()
    lat.step(40)
    d = lat.diagnostics(slices=True)
    mac = lat.calmacros()
    live = np.ones(mac["rho"].shape, bool)
    _check(d, _host_sums(mac, live), d["mass"])
    area = ((1.0 + mac["phi"]) / 2.0).sum(axis=(0, 1))
    assert np.max(np.abs(d["slice_red"] - area)) <= 1e-12 * area.max()
    tol = 1e-13 if precision == "f64" else 1e-6
    assert abs(d["mass"] / d0["mass"] - 1.0) < tol
    assert abs(d["jz"]) < 1e-9 * d["mass"]  # symmetric jet: no net momentum


def test_diagnostics_walls_and_particles():
    cfg = ic.particle_config2(n=32, rad=5.0)
    parts = ic.sphere_particles([(16.3, 15.8, 16.4)], 5.0, cfg["rho_p"], cfg["cita_deg"], cfg["sigma"])
    ri, bi, _ = ic.flat_interface_state(32, 32, 32, z0=16.0, rho0=1.0)
    lat = symmetry_amd.Lattice(32, 32, 32, precision="f64", sigma=cfg["sigma"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"],
                               bc=cfg["bc"], npar=1)
    lat.upload(ri, bi)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"],
                      up=np.array([[0.05, 0.0, 0.0]]))
    lat.step(6)
    d = lat.diagnostics(slices=True)
    mac = lat.calmacros()
    masks = lat.masks()
    live = masks["walls_all"] == 0
    live[:, :, 0] = live[:, :, -1] = False  # z wall ghost layers
    _check(d, _host_sums(mac, live), d["mass"])
    area = np.where(live, (1.0 + mac["phi"]) / 2.0, 0.0).sum(axis=(0, 1))
    assert np.max(np.abs(d["slice_red"] - area)) <= 1e-12 * area.max()
