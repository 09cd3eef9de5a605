"""Pins of the oracle's lattice tables and equilibria by closed forms and by
the reference's own DATA tables (tests/golden/d3q19_tables.json, parsed from
Source1.f90 S:105-142 by tests/golden/make_tables_fixture.py)."""
import json
import math
import os

import numpy as np

import oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "d3q19_tables.json")


def _gold():
    with open(GOLD) as f:
        return json.load(f)


def test_tables_match_reference_data():
    d = _gold()
    wi, c, opc, sym, M, Minv = oracle.tables()
    assert np.array_equal(M, np.array(d["M"], dtype=float))
    assert np.array_equal(c, np.array([d["cix"], d["ciy"], d["ciz"]]))
    assert np.array_equal(opc, d["opc"])
    assert np.array_equal(sym, np.array([d["symx"], d["symy"], d["symz"]]))
    assert np.array_equal(wi, 1.0 / np.array(d["wi_denominator"], dtype=float))


def test_moment_matrix_closed_forms():
    wi, c, opc, sym, M, Minv = oracle.tables()
    G = M @ M.T
    norms = np.diag(G)
    # Rows are orthogonal, with the norms M M^T of the reference's matrix (S:105-123).
    assert np.allclose(G - np.diag(norms), 0.0)
    assert list(norms) == [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8]
    assert np.count_nonzero(M) == 213
    assert np.array_equal(M[0], np.ones(19))
    assert np.array_equal(M[3], c[0]) and np.array_equal(M[5], c[1]) and np.array_equal(M[7], c[2])
    assert np.allclose(Minv @ M, np.eye(19), atol=1e-14)
    # weights: isotropy up to 4th order
    assert math.isclose(wi.sum(), 1.0, rel_tol=0, abs_tol=1e-15)
    for a in range(3):
        assert abs(np.sum(wi * c[a])) < 1e-16
        for b in range(3):
            assert math.isclose(np.sum(wi * c[a] * c[b]), (1.0 / 3.0) * (a == b), abs_tol=1e-16)


def test_opposite_and_mirror_maps():
    wi, c, opc, sym, M, Minv = oracle.tables()
    for i in range(19):
        assert np.array_equal(c[:, opc[i]], -c[:, i])
        for a in range(3):
            e = c[:, i].copy()
            e[a] = -e[a]
            assert np.array_equal(c[:, sym[a, i]], e)


def test_equilibrium_moments_identity():
    """meq of equilf (S:405-426) is exactly M @ equilf_SRT (S:841-845)."""
    wi, c, opc, sym, M, Minv = oracle.tables()
    rng = np.random.default_rng(3)
    for _ in range(20):
        rho = rng.uniform(0.5, 1.5)
        u = rng.uniform(-0.1, 0.1, 3)
        f = rng.uniform(-1e-3, 1e-3, 3)
        meq, fk = oracle.equilf(rho, *u, *f)
        feq = oracle.equilf_srt(rho, *u)
        assert np.allclose(M @ feq, meq, rtol=1e-14, atol=1e-15)
        # Guo source populations w_i[3(c-u).F + 9(c.u)(c.F)] have moments fk.
        cu = c.T @ u
        cf = c.T @ f
        guo = wi * (3.0 * (cf - u @ f) + 9.0 * cu * cf)
        assert np.allclose(M @ guo, fk, rtol=0, atol=1e-15)


def test_lagrange_point_count():
    pi = 4.0 * math.atan(1.0)
    alpha = 2.0 / 180.0 * pi
    la_m, la_n = int(2.0 * pi / alpha), int(pi / alpha)
    assert (la_m, la_n, la_m * (la_n + 1)) == (180, 90, 16380)
