"""Pins of the oracle where the reference has no counterpart: its geometry has
no periodic faces and only low-side walls (x=1, z=1; S:806-807, S:1557-1590),
so the periodic wrap and the high-side walls are pinned by symmetry.

* Periodic translation: every node runs the same operations in the same
  order on its own neighbourhood, so stepping a translated state gives the
  translated result bit for bit.  A wrong wrap index breaks it.
* Reflection between two walls: mirroring the state in z (nodes z -> nz-1-z,
  directions i -> sym_z(i)) and stepping gives the mirrored result; only the
  summation order of direction sums changes, so agreement is to round-off.
  A wrong high-side ghost plane or bounce-back direction breaks it at O(1).
"""
import numpy as np

import oracle
from symmetry_amd import ic


def _state(nx, ny, nz, bc, **kw):
    st = oracle.OracleState(nx, ny, nz, bc=bc, sigma=0.01, rho0=1.0, nu_r=0.1, nu_b=1.0 / 6.0, beta=0.7)
    return st


def _drop_state(nx, ny, nz, center, r0):
    ri, bi, _ = ic.flat_interface_state(nx, ny, nz, z0=-1e3, rho0=1.0, drop=(*center, r0), sigma=0.01)
    return ri, bi


def test_periodic_translation_invariance():
    nx, ny, nz = 12, 10, 16
    ri, bi = _drop_state(nx, ny, nz, (5.0, 4.5, 7.0), 3.5)
    warm = _state(nx, ny, nz, (0,) * 6)
    warm.ri[...] = ri
    warm.bi[...] = bi
    warm.step(4)  # develop a flow so every stage has non-trivial input
    r0, b0 = warm.ri.copy(), warm.bi.copy()
    shift = (3, 7, 5)
    a = _state(nx, ny, nz, (0,) * 6)
    a.ri[...] = r0
    a.bi[...] = b0
    a.step(3)
    b = _state(nx, ny, nz, (0,) * 6)
    b.ri[...] = np.roll(r0, shift, axis=(0, 1, 2))
    b.bi[...] = np.roll(b0, shift, axis=(0, 1, 2))
    b.step(3)
    assert np.array_equal(b.ri, np.roll(a.ri, shift, axis=(0, 1, 2)))
    assert np.array_equal(b.bi, np.roll(a.bi, shift, axis=(0, 1, 2)))


def _mirror_z(pops, sym):
    return pops[:, :, ::-1, :][..., sym[2]]


def test_two_wall_reflection_symmetry():
    nx, ny, nz = 10, 8, 14
    bc = (0, 0, 0, 0, 1, 1)
    _, _, _, sym, _, _ = oracle.tables()
    ri, bi = _drop_state(nx, ny, nz, (4.5, 4.0, 3.5), 3.0)  # touching the z=1 wall
    a = _state(nx, ny, nz, bc)
    a.ri[...] = ri
    a.bi[...] = bi
    a.step(6)
    b = _state(nx, ny, nz, bc)
    b.ri[...] = _mirror_z(ri, sym)
    b.bi[...] = _mirror_z(bi, sym)
    b.step(6)
    live = (slice(None), slice(None), slice(1, nz - 1))
    ra, rb = _mirror_z(a.ri, sym)[live], b.ri[live]
    assert np.max(np.abs(ra - rb)) < 1e-13
    assert np.max(np.abs(_mirror_z(a.bi, sym)[live] - b.bi[live])) < 1e-13
    # the flow is non-trivial and reaches the wall
    a.call("calmacros")
    assert np.max(np.abs(a.uz[:, :, 1])) > 1e-6


def test_reflection_test_detects_a_wrong_wall():
    """The reflection check fails when one wall is not a wall: with the high
    z face periodic the mirrored runs part at O(1e-4) or more."""
    nx, ny, nz = 10, 8, 14
    _, _, _, sym, _, _ = oracle.tables()
    ri, bi = _drop_state(nx, ny, nz, (4.5, 4.0, 3.5), 3.0)
    a = _state(nx, ny, nz, (0, 0, 0, 0, 1, 1))
    a.ri[...] = ri
    a.bi[...] = bi
    a.step(6)
    b = _state(nx, ny, nz, (0, 0, 0, 0, 1, 0))  # planted mistake: z-high face not a wall
    b.bi[...] = _mirror_z(bi, sym)
    b.ri[...] = _mirror_z(ri, sym)
    b.step(6)
    live = (slice(None), slice(None), slice(1, nz - 1))
    assert np.max(np.abs(_mirror_z(a.ri, sym)[live] - b.ri[live])) > 1e-6
