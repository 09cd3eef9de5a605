"""GPU: cgl_diagnostics_noalloc, the allocation-free variant of cgl_diagnostics
(device sums of mass, colour masses, momentum and kinetic energy over the
live fluid nodes, and the red area of every z slice; the mass check and
Energy.dat sums of S:1673-1797), called through the C ABI.

Its kernels form each node's moments with calmacros' own code and reduce
into a static device buffer, in another fixed order than cgl_diagnostics'
field buffer and block tree: the results agree to rounding (1e-12 relative
to the mass for the sums, 1e-12 relative for the slice areas).  Each case
calls both entry points on two handles in a row, the second one smaller,
so the shared static buffer is reused with fewer partials; a box too large
for the buffer is rejected with an error before any kernel runs.
"""
import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import small_rp

pytestmark = pytest.mark.gpu

WALLS_Z = (0, 0, 0, 0, symmetry_amd.BC_WALL, symmetry_amd.BC_WALL)


def _print(*a):
    print(*a, flush=True)


def _fluid(shape, precision, layout, bc=(0,) * 6):
    nx, ny, nz = shape
    cfg, ri, bi = small_rp(nx=nx, ny=ny, nz=nz, r0=min(nx, ny) / 4.0, nu_b=0.1)
    lat = symmetry_amd.Lattice(nx, ny, nz, precision=precision, sigma=cfg["sigma"], rho0=cfg["rho0"],
                               nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], layout=layout, bc=bc)
    if bc == (0,) * 6:
        lat.upload(ri, bi)
    else:
        lat.init_interface(nz / 2.0, 6.0 * cfg["beta"] * 0.134)
    lat.step(5)
    return lat


def _particles(n, precision):
    c2 = ic.particle_config2(n=n, rad=n / 8.0)
    parts = ic.sphere_particles([c2["center"]], c2["rad"], c2["rho_p"], c2["cita_deg"], c2["sigma"])
    lat = symmetry_amd.Lattice(n, n, n, precision=precision, sigma=c2["sigma"], rho0=c2["rho0"], nu_r=c2["nu_r"],
                               nu_b=c2["nu_b"], beta=c2["beta"], bc=c2["bc"], npar=1)
    lat.init_interface(c2["z0"], 6.0 * c2["beta"] * 0.134)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"],
                      up=[[0.01, 0.0, -0.02]])
    lat.step(4)
    return lat


def _compare(lat, tag):
    ref = lat.diagnostics(slices=True)
    got = lat.diagnostics(slices=True, noalloc=True)
    mass = abs(ref["mass"])
    for k in ("mass", "mass_r", "mass_b", "jx", "jy", "jz", "ekin"):
        assert abs(got[k] - ref[k]) <= 1e-12 * mass + 1e-12 * abs(ref[k]), (tag, k, got[k], ref[k])
    np.testing.assert_allclose(got["slice_red"], ref["slice_red"], rtol=1e-12, atol=1e-12 * np.max(ref["slice_red"]))
    _print(tag, "mass", got["mass"], "ekin", got["ekin"], "max slice diff",
           float(np.max(np.abs(got["slice_red"] - ref["slice_red"]))))


CASES = [
    ("f64-rb periodic", lambda s: _fluid(s, "f64", "rb")),
    ("f32-fs scatter", lambda s: _fluid(s, "f32", "fs")),
    ("f16-fs scatter", lambda s: _fluid(s, "f16", "fs")),
    ("f64-fs periodic", lambda s: _fluid(s, "f64", "fs")),
    ("f32-rb z walls", lambda s: _fluid(s, "f32", "rb", WALLS_Z)),
]


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_noalloc_diagnostics_match(name, make):
    for shape in ((40, 24, 128), (20, 12, 64)):  # the second call: a smaller box
        _print(name, shape)
        lat = make(shape)
        _compare(lat, f"{name} {shape}")
        lat.close()


def test_noalloc_diagnostics_with_particles():
    for n in (48, 32):
        _print("particles f64", n)
        lat = _particles(n, "f64")
        _compare(lat, f"particles {n}")
        lat.close()


def test_noalloc_rejects_boxes_beyond_the_buffer():
    # 8 + nz + 7 gz gy + gy nz doubles: gy = 1024, nz = 512 needs more than 2^19
    lat = symmetry_amd.Lattice(1, 4096, 512, precision="f32", layout="rb")
    with pytest.raises(symmetry_amd.CglError, match="static buffer"):
        lat.diagnostics(noalloc=True)
    lat.close()
