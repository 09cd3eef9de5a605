"""The oracle's contact-angle geometry against closed forms, for angles other
than the reference's compile-time 90 degrees (S:21).

Two constructions of the reference use the contact angle c:

* the wetting correction of the interface normal at particle-adjacent fluid
  nodes (S:536-587): the normal n0 is replaced by the unit vector at angle c
  from the particle's outward normal nw, in the plane of nw and n0, on the
  side of n0; the gradient keeps its length and takes that direction;
* the capillary force on the Lagrange points (S:1995-2022): with ln the
  corrected normal at the point, ie = -ln, the force direction is
  unit((ie x nw) x ie) and its size carries |ie x nw|.

Written in the basis (nw, t), t = unit(n0 - (n0 . nw) nw), the corrected
normal is ln = cos(c) nw + sin(c) t, the capillary direction is
sin(c) nw - cos(c) t (the unit vector in that plane perpendicular to ln, on
the side of nw) and |ie x nw| = sin(c).  The checks below evaluate these
forms directly, on fields where every input of the oracle is known: a linear
phase field (whose D3Q19 gradient is exact) for the wetting pass, and a
uniform node normal with phi = 0 at every node for the capillary force.  A
restatement that takes the other root of the angle condition, drops the
projection factor or flips the tangent fails them; each test shows that the
mirrored construction differs from the oracle by far more than its
tolerance.
"""
import math

import numpy as np
import pytest

import oracle

PI = 4.0 * math.atan(1.0)
KCAP = 0.134 * 1.43  # S:1989
ALPHA = 2.0 / 180.0 * PI  # S:22


def _unit(v):
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _tangent(nw, n0):
    """t: the unit vector in the plane of nw and n0, perpendicular to nw, on
    the side of n0."""
    return _unit(n0 - np.sum(n0 * nw, axis=-1, keepdims=True) * nw)


N0 = _unit(np.array([0.31, -0.52, 0.79]))


@pytest.mark.parametrize("deg", [30.0, 60.0, 90.0, 120.0, 150.0])
def test_wetting_normal_closed_form(deg):
    n = 32
    center = np.array([16.3, 15.7, 16.45])
    rad = 6.0
    st = oracle.OracleState(n, n, n, npar=1)
    st.center[0] = center
    st.rad[0] = rad
    st.cita[0] = deg / 180.0 * PI
    g0 = 0.01 * N0
    x, y, z = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    st.phi[...] = g0[0] * x + g0[1] * y + g0[2] * z
    # fluid nodes within one lattice unit outside the sphere (1-based coordinates)
    pos = np.stack([x + 1.0, y + 1.0, z + 1.0], -1)
    d = np.linalg.norm(pos - center, axis=-1)
    shell = (d > rad) & (d < rad + 1.0)
    st.pout[...] = np.where(shell, 1, 0)
    st.call("derivatives")

    nw = _unit(pos[shell] - center)
    t = _tangent(nw, np.broadcast_to(N0, nw.shape))
    c, s = math.cos(st.cita[0]), math.sin(st.cita[0])
    want = c * nw + s * t
    got = np.stack([st.nxv[shell], st.nyv[shell], st.nzv[shell]], -1)
    assert shell.sum() > 100
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-9)
    grad = np.stack([st.gx[shell], st.gy[shell], st.gz[shell]], -1)
    np.testing.assert_allclose(grad, np.linalg.norm(g0) * want, rtol=0, atol=1e-11)
    # nodes off the shell keep the exact gradient direction n0
    inner = (d > rad + 2.0) & (d < rad + 6.0)
    np.testing.assert_allclose(st.nxv[inner], N0[0], rtol=0, atol=1e-12)
    # the mirrored root (the side away from n0) is far from what the oracle gives
    if deg != 90.0:
        assert np.max(np.abs(got - (c * nw - s * t))) > 0.1


def _lagrange_points(center, rad):
    """The Lagrange points and surface elements of S:1678-1690."""
    la_m, la_n = int(2.0 * PI / ALPHA), int(PI / ALPHA)
    aa = np.arange(la_m)[:, None] * ALPHA
    bb = np.arange(la_n + 1)[None, :] * ALPHA
    pts = np.stack(np.broadcast_arrays(center[0] + rad * np.sin(bb) * np.cos(aa),
                                       center[1] + rad * np.sin(bb) * np.sin(aa),
                                       center[2] + rad * np.cos(bb)), -1).reshape(-1, 3)
    ds = np.broadcast_to(rad * rad * np.sin(bb) * ALPHA * ALPHA, (la_m, la_n + 1)).reshape(-1)
    return pts, ds


@pytest.mark.parametrize("deg", [45.0, 60.0, 90.0, 135.0])
def test_capillary_force_closed_form(deg):
    n = 32
    center = np.array([16.3, 15.7, 16.45])
    rad = 7.5
    sigma, beta = 0.05, 0.7
    st = oracle.OracleState(n, n, n, npar=1, sigma=sigma, beta=beta)
    st.center[0] = center
    st.rad[0] = rad
    cita = deg / 180.0 * PI
    st.cita[0] = cita
    st.nxv[...], st.nyv[...], st.nzv[...] = N0
    st.call("lagrange_capillary")

    pts, ds = _lagrange_points(center, rad)
    nw = _unit(pts - center)
    t = _tangent(nw, np.broadcast_to(N0, nw.shape))
    c, s = math.cos(cita), math.sin(cita)
    star = 4.5 * KCAP * beta * ds * s  # phi = 0 at every point: (1 - phi^2)^2 = 1
    direction = nw if deg == 90.0 else s * nw - c * t
    f = (star * sigma)[:, None] * direction
    want_f = f.sum(0)
    want_t = np.cross(pts - center, f).sum(0)
    scale = np.sum(star * sigma)
    print("tension", st.tension[0], "want", want_f, "scale", scale)
    np.testing.assert_allclose(st.tension[0], want_f, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(st.tension_t[0], want_t, rtol=0, atol=1e-12 * scale * rad)
    assert st.shb[0] == 0.0  # no point sees phi > 0
    if deg != 90.0:
        # the tangent follows the uniform normal, so the sum over the sphere
        # is not zero (at 90 degrees the force is radial and sums to zero)
        assert np.linalg.norm(want_f) > 1e-3 * scale
        wrong = (star * sigma)[:, None] * (s * nw + c * t)
        assert np.linalg.norm(wrong.sum(0) - want_f) > 1e-2 * scale
