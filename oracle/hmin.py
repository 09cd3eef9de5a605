"""Thinnest-thread record of output3 (Source1.f90 S:1673-1717), restated.

TEST INFRASTRUCTURE ONLY (part of the oracle package): the checker of the
hmin.dat lines the Fortran driver writes.  Pinned against the compiled
reference's own output3 (tests/test_oracle_vs_ref.py).

Arrays are indexed [x, y, z] with 0-based indices (the reference's (z, y, x)
Fortran arrays read in C order).  On the plane x = xl - 1 (1-based), for each
y = 2..yl-1 the reference scans z = 2..zl-1 upwards and, wherever the node
and the node above are both fluid (pout = 0 and walls_all = 0) and rho_N
changes sign between them, sets bottom(y) = z+1 - rho_N(z+1) / (rho_N(z+1) -
rho_N(z)): the last such crossing counts, and a column without one keeps
bottom = 1 (S:1681-1697).  h(y) = (zl - 1/2) - bottom(y); hmin is the first
smallest h over y, y0 its y (S:1699-1705).  Where no h is below zl the
reference leaves y0 undefined; this restatement, like the Fortran driver,
reports 0 there.
"""
from __future__ import annotations


def hmin(phi, walls_all, pout):
    """(hmin, y0) of output3 for fields indexed [x, y, z]; y0 is 1-based."""
    nx, ny, nz = phi.shape
    x = nx - 2  # the plane x = xl - 1
    best, y0 = float(nz), 0
    for y in range(1, ny - 1):
        bottom = 1.0
        col, wal, pou = phi[x, y], walls_all[x, y], pout[x, y]
        for z in range(1, nz - 1):
            zn = z + 1
            if pou[z] == 0 and wal[z] == 0 and pou[zn] == 0 and wal[zn] == 0:
                a, b = float(col[z]), float(col[zn])
                if a * b < 0.0:
                    bottom = float(zn + 1) - b / (b - a)
        h = (float(nz) - 0.5) - bottom
        if h < best:
            best, y0 = h, y + 1
    return best, y0


def line(t, tr, h, y0, rad_c, lamda_d):
    """The hmin.dat record (format (1(1xI8),3(1xF12.6)), S:1713)."""
    return " %8d" % t + "".join(" %12.6f" % v for v in (t / tr, h / rad_c, y0 / lamda_d))
