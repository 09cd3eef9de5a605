"""Energy budget of output4 (Source1.f90 S:1719-1797), restated in numpy.

TEST INFRASTRUCTURE ONLY (part of the oracle package): the checker of the
Energy.dat lines the Fortran driver writes.  Pinned against the compiled
reference's own output4 (tests/test_oracle_vs_ref.py).

Arrays are indexed [x, y, z] with 0-based indices (the reference's (z, y, x)
Fortran arrays read in C order).  The sums run over the nodes x, y, z =
2..n-1 (1-based) whose ``walls_all`` is <= 0, and the reference reads
``walls_all``, ``rho`` and ``uz`` there at x index = z (``walls_all(z,y,z)``,
``rho(z,y,z)``, ``uz(z,y,z)``, S:1745-1750) while ``rho_N``, ``ux`` and ``uy``
are read at the node itself; this restatement reads them the same way where
z <= nx (1-based; always in the reference's cube) and, like the Fortran
driver, at the node itself where z > nx, which the reference's expression
does not define.
"""
from __future__ import annotations

import math

import numpy as np

PI = 4.0 * math.atan(1.0)  # S:20
NEAR = float(np.float32(0.1))  # the single-precision literal 0.1 of the plane tests (S:1779-1791)

# the 18 moving D3Q19 directions (the area test asks for any of them)
_C = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1),
      (1, 1, 0), (-1, -1, 0), (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1),
      (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1), (0, -1, 1)]


def field_terms(rho_n, rho, ux, uy, uz, walls_all, rho0):
    """(Epv, Emov, area) before scaling (S:1736-1762): pressure term and
    kinetic energy over red live nodes, and the count of live nodes with a
    live neighbour of the other colour."""
    nx, ny, nz = rho_n.shape
    sx, sy, sz = slice(1, nx - 1), slice(1, ny - 1), slice(1, nz - 1)
    shape = (nx - 2, ny - 2, nz - 2)
    # values read at x index = z (0-based z < nx), else at the node
    zq = min(nz - 1, nx)  # interior z of 1..zq-1 take the (z, y, z) reads
    y = np.arange(1, ny - 1)[:, None]
    z = np.arange(1, zq)[None, :]

    def quirk(a):
        out = np.array(np.broadcast_to(a[sx, sy, sz], shape))
        out[:, :, : zq - 1] = a[z, y, z][None, :, :]
        return out

    live = quirk(walls_all) <= 0
    rho_zyz = quirk(rho)
    uz_zyz = quirk(uz)
    phi = rho_n[sx, sy, sz]
    red = live & (phi > 0.0)
    epv = np.sum(np.where(red, (rho[sx, sy, sz] - rho0) / 3.0, 0.0))
    emov = np.sum(np.where(red, 0.5 * rho_zyz * (ux[sx, sy, sz] ** 2 + uy[sx, sy, sz] ** 2 + uz_zyz ** 2), 0.0))
    hit = np.zeros(phi.shape, dtype=bool)
    for cx, cy, cz in _C:
        nb = (slice(1 + cx, nx - 1 + cx), slice(1 + cy, ny - 1 + cy), slice(1 + cz, nz - 1 + cz))
        hit |= (phi * rho_n[nb] < 0.0) & (walls_all[nb] <= 0)
    area = float(np.count_nonzero(live & hit))
    return epv, emov, area


def energy(rho_n, rho, ux, uy, uz, walls_all, *, rho0, sigma, rad_c, up, rp, tension, tension_t, mass, inertia,
           center0, shb, wcap_old):
    """output4's quantities (S:1719-1797).  Particle arrays are (par, 3)
    (up, rp, tension, tension_t, center0) or (par,) (mass, inertia, shb);
    the reference passes one mass and one moment of inertia for all
    particles.  Returns a dict with the eight Energy.dat columns after t and
    t*, the scaled Shb per particle and the new Wcap_old."""
    nx, ny, nz = rho_n.shape
    er = PI * rad_c ** 2 * sigma
    epv, emov, area = field_terms(rho_n, rho, ux, uy, uz, walls_all, rho0)
    epv = epv / er * 8.0
    emov = emov / er * 8.0
    esur = area / 2.0 * sigma / er * 8.0
    npar = len(shb)
    emov_p = [0.0] * npar
    wcap = [0.0] * npar
    shb = [float(v) for v in shb]
    for p in range(npar):
        u, w = up[p], rp[p]
        emov_p[p] = (0.5 * mass[p] * (u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
                     + 0.5 * inertia[p] * (w[0] ** 2 + w[1] ** 2 + w[2] ** 2))
        wcap[p] = (tension[p][0] * u[0] + tension[p][1] * u[1] + tension[p][2] * u[2]
                   + tension_t[p][0] * w[0] + tension_t[p][1] * w[1] + tension_t[p][2] * w[2])
        c0 = center0[p]
        # particles set up on a symmetry plane count half (S:1778-1792)
        if abs(c0[1] - (ny - 0.5)) < NEAR or abs(c0[1] - 1.5) < NEAR:
            emov_p[p] *= 0.5
            wcap[p] *= 0.5
            shb[p] *= 0.5
        if abs(c0[0] - (nx - 0.5)) < NEAR:
            emov_p[p] *= 0.5
            wcap[p] *= 0.5
            shb[p] *= 0.5
        if abs(c0[2] - (nz - 0.5)) < NEAR:
            emov_p[p] *= 0.5
            wcap[p] *= 0.5
            shb[p] *= 0.5
    s_emov, s_wcap = 0.0, 0.0
    for p in range(npar):  # sequential sums, as Fortran's SUM
        s_emov += emov_p[p]
        s_wcap += wcap[p]
    emov_psum = s_emov / er * 8.0
    wcap_sum = wcap_old + s_wcap / er * 8.0
    shb = [v / (PI * rad_c * rad_c) * 8.0 for v in shb]
    s_shb = 0.0
    for v in shb:
        s_shb += v
    return {"Esur": esur, "Shb": s_shb, "Efree": esur - epv, "Emov": emov, "Wcap": wcap_sum, "Emov_p": emov_psum,
            "Esum": esur - epv + emov + emov_psum, "shb": np.array(shb), "wcap_old": wcap_sum}


COLUMNS = ("Esur", "Shb", "Efree", "Emov", "Wcap", "Emov_p", "Esum")


def line(t, tr, e):
    """The Energy.dat record (format (1(1xI8),8(1xF12.6)), S:1796)."""
    return " %8d" % t + "".join(" %12.6f" % v for v in [t / tr] + [e[k] for k in COLUMNS])
