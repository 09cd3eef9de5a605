"""The oracle against the unmodified reference, stage by stage.

The reference (Source1.f90) is compiled from the reference checkout into
oracle/_ref/libref.so by `make -C oracle ref` and driven through
oracle/ref_harness.f90 at its own compile-time box: 152^3, 12 particles, the
1/8-symmetric geometry (walls at x=1 and z=1, mirrors elsewhere).  Every
field, mask and particle quantity must agree bit for bit: both sides run the
same FP64 operations in the same order with FMA contraction off.

Skipped where the reference build is absent (e.g. a checkout without the
reference).
"""
import json
import math
import os

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref/libref.so not built")

GOLD = os.path.join(os.path.dirname(__file__), "golden", "d3q19_tables.json")

# Reference parameters (S:58-84).
OH, LAM, RAD_C, SIGMA, RHO0, RAD = 0.1, 10.0, 30.0, 0.05, 0.05, 7.5
NU_R = OH / math.sqrt(RHO0 / (RAD_C * SIGMA))
NU_B = NU_R / LAM
PI = 4.0 * math.atan(1.0)
MASS = 4.0 / 3.0 * RHO0 * PI * RAD ** 3
INERTIA = 2.0 / 5.0 * MASS * RAD ** 2
FG = -PI * 2.0 * RAD * SIGMA

PV = {"fp": ("Fpx", "Fpy", "Fpz"), "tp": ("Tpx", "Tpy", "Tpz"), "frep": ("Frex", "Frey", "Frez"),
      "up": ("upx", "upy", "upz"), "rp": ("rpx", "rpy", "rpz")}


@pytest.fixture(scope="module")
def ref():
    with open(GOLD) as f:
        M = np.array(json.load(f)["M"], dtype=float)
    h = oracle.RefHarness(M, SIGMA, RHO0, NU_R, NU_B, RAD, RAD_C)
    sc = h.scal()
    sc[0], sc[1], sc[2], sc[3] = FG, MASS, INERTIA, (h.yl - 2) * 2.0
    return h


def make_oracle(ref):
    xl = ref.xl
    st = oracle.OracleState(xl, ref.yl, ref.zl, bc=oracle.REF_BC, npar=ref.par, sigma=SIGMA, rho0=RHO0,
                            nu_r=NU_R, nu_b=NU_B, fg=FG)
    s = st.s
    # repulsion planes and Lagrange in-domain box of the reference (S:1607-1612, S:1832)
    for a, (lo, hi) in enumerate([(1.5, xl - 1.5), (1.5, ref.yl - 1.5), (1.5, ref.zl - 0.5)]):
        s.rep_plane[2 * a], s.rep_plane[2 * a + 1] = lo, hi
        s.rep_on[2 * a] = s.rep_on[2 * a + 1] = 1
    s.dom_hi[0], s.dom_hi_on[0] = xl - 0.5, 1
    s.dom_lo[1], s.dom_lo_on[1] = 1.5, 1
    s.dom_hi[1], s.dom_hi_on[1] = ref.yl - 0.5, 1
    s.dom_hi[2], s.dom_hi_on[2] = ref.zl - 0.5, 1
    s.ref_rules = 1
    st.rad[:] = RAD
    st.mass[:] = MASS
    st.inertia[:] = INERTIA
    st.cita[:] = 90.0 / 180.0 * PI
    return st


def copy_ref_to_oracle(ref, st):
    for k in oracle.REF_FIELDS:
        st.arrays[k][...] = ref.field(k)
    for k in ("center", "center_old", "center0", "tension", "tension_t"):
        st.arrays[k][...] = ref.part(k)
    for k, names in PV.items():
        st.arrays[k][...] = np.stack([ref.part(n) for n in names], 1)
    st.shb[:] = ref.part("Shb")


def mismatches(ref, st):
    bad = []
    for k in oracle.REF_FIELDS:
        a, b = st.arrays[k], ref.field(k)
        if not np.array_equal(a, b):
            bad.append(k)
    for k in ("center", "center_old", "tension", "tension_t"):
        if not np.array_equal(st.arrays[k], ref.part(k)):
            bad.append(k)
    for k, names in PV.items():
        if not np.array_equal(st.arrays[k], np.stack([ref.part(n) for n in names], 1)):
            bad.append(k)
    if not np.array_equal(st.shb, ref.part("Shb")):
        bad.append("shb")
    return bad


def test_fluid_stages_bit_exact(ref):
    """calmacros, derivatives, collision, Recolor, boundary, streaming, curve
    with the particles parked outside the box (no masks)."""
    ref.part("center")[:] = -1000.0
    ref.call("set_walls")
    ref.call("initialization")
    st = make_oracle(ref)
    copy_ref_to_oracle(ref, st)
    for rstage, ostage in [("calmacros", "calmacros"), ("derivatives", "derivatives"), ("collision", "collision"),
                           ("recolor", "recolor"), ("boundary", "boundary"), ("streaming_r", "ri"),
                           ("streaming_b", "bi"), ("curve", "curve")]:
        ref.call(rstage)
        if rstage.startswith("streaming"):
            st.streaming(ostage)
        else:
            st.call(ostage)
        assert mismatches(ref, st) == [], rstage
    assert np.count_nonzero(ref.field("flag")) > 0  # the interface is present


@pytest.mark.slow
def test_particle_steps_bit_exact(ref):
    """Two full steps with the reference's 12 particles (S:209-214 start-up,
    then S:216-251), particles pushed so that nodes get uncovered and refilled,
    two of them moved into pair-repulsion range."""
    ref.call("initia_center")
    ref.call("set_walls")
    ref.call("initialization")
    ref.call("derivatives")
    # drive the particles so refilling (S:1351-1418) has work
    ref.part("upx")[:] = np.linspace(-0.6, 0.6, ref.par)
    ref.part("upy")[:] = 0.45
    ref.part("rpz")[:] = 0.01
    st = make_oracle(ref)
    copy_ref_to_oracle(ref, st)
    assert mismatches(ref, st) == []
    order = [("calmacros", "calmacros"), ("lagrange", "lagrange_capillary"), ("capillary", None),
             ("momentum", "momentum"), ("repulsive", "repulsive"), ("movement", "movement"),
             ("set_walls", "set_walls"), ("refilling", "refilling"), ("derivatives", "derivatives"),
             ("collision", "collision"), ("recolor", "recolor"), ("boundary", "boundary"),
             ("streaming_r", "ri"), ("streaming_b", "bi"), ("curve", "curve")]
    refilled = 0
    for step in range(2):
        for rstage, ostage in order:
            if rstage == "refilling":
                refilled += int(np.count_nonzero((ref.field("pout") >= 1) & (ref.field("walls_old") >= 1)))
            ref.call(rstage)
            if ostage is None:
                pass
            elif rstage.startswith("streaming"):
                st.streaming(ostage)
            else:
                st.call(ostage)
            if rstage != "lagrange":
                assert mismatches(ref, st) == [], (step, rstage)
    assert refilled > 0
    assert np.count_nonzero(ref.field("pout")) > 0 and np.count_nonzero(ref.field("pin")) > 0


def test_pair_repulsion_bit_exact(ref):
    ref.call("initia_center")
    c = ref.part("center")
    c[1] = c[0] + np.array([0.0, 2 * RAD + 1.2, 0.0])  # gap 1.2 < range 3
    st = make_oracle(ref)
    copy_ref_to_oracle(ref, st)
    ref.call("repulsive")
    st.call("repulsive")
    assert mismatches(ref, st) == []
    assert np.any(ref.part("Frey") != 0.0)


def test_output4_energy_line(ref):
    """The reference's own output4 (S:1719-1797) at its box with its 12
    particles, which initia_center places on the mirror planes (so the
    half-count rules act), with seeded rho, u, particle velocities, capillary
    forces, Shb and Wcap_old: its Energy.dat line equals the restatement's
    line (format (1(1xI8),8(1xF12.6))), and Wcap_old and Shb, which output4
    updates in place, agree bit for bit."""
    from oracle import energy as en

    ref.call("initia_center")
    ref.call("set_walls")
    ref.call("initialization")
    rng = np.random.default_rng(5)
    shape = ref.field("rho").shape
    ref.field("rho")[...] = 1.0 + 0.01 * rng.standard_normal(shape)
    for k in ("ux", "uy", "uz"):
        ref.field(k)[...] = 0.01 * rng.standard_normal(shape)
    for names in (PV["up"], PV["rp"]):
        for n in names:
            ref.part(n)[:] = 0.05 * rng.standard_normal(ref.par)
    ref.part("tension")[...] = rng.standard_normal((ref.par, 3))
    ref.part("tension_t")[...] = rng.standard_normal((ref.par, 3))
    ref.part("Shb")[:] = rng.uniform(10.0, 100.0, ref.par)
    tr, t, wcap_old = 123.25, 400, 0.3
    sc = ref.scal()
    sc[4], sc[5], sc[6] = wcap_old, tr, float(t)
    c0 = ref.part("center0").copy()
    assert np.any(np.abs(c0[:, 1] - 1.5) < 0.1) and np.any(np.abs(c0[:, 0] - (ref.xl - 0.5)) < 0.1)
    e = en.energy(ref.field("phi").copy(), ref.field("rho").copy(), ref.field("ux").copy(), ref.field("uy").copy(),
                  ref.field("uz").copy(), ref.field("walls_all").copy(), rho0=RHO0, sigma=SIGMA, rad_c=RAD_C,
                  up=np.stack([ref.part(n) for n in PV["up"]], 1), rp=np.stack([ref.part(n) for n in PV["rp"]], 1),
                  tension=ref.part("tension").copy(), tension_t=ref.part("tension_t").copy(),
                  mass=[MASS] * ref.par, inertia=[INERTIA] * ref.par, center0=c0, shb=ref.part("Shb").copy(),
                  wcap_old=wcap_old)
    assert np.count_nonzero(ref.field("walls_all")) > 0 and e["Esur"] > 0 and e["Emov"] > 0
    ref.call("output4")
    assert ref.energy_line() == en.line(t, tr, e)
    assert sc[4] == e["wcap_old"]
    assert np.array_equal(ref.part("Shb"), e["shb"])


def test_output3_hmin_line(ref):
    """The reference's own output3 (S:1673-1717) on designed fields of its
    box, against the restatement's hmin.dat line.  On the plane x = xl - 1
    an interface whose height varies with y (its lowest column, the neck,
    is not the first), a second, lower sign change in some columns (the last
    crossing counts), crossings next to a wall or pout node in other columns
    (skipped), and a few columns without a crossing (bottom stays 1); the
    neighbouring planes hold other heights, so reading the wrong plane, the
    first crossing, or z for z+1 changes the line."""
    from oracle import hmin as hm

    xl, yl, zl = ref.xl, ref.yl, ref.zl
    sc = ref.scal()
    lamda_d = (yl - 2) * 2.0
    sc[3] = lamda_d
    rng = np.random.default_rng(11)
    zz = np.arange(zl, dtype=float) + 1.0
    for t, tr in ((0, 123.25), (400, 123.25), (1200, 77.5)):
        phi, wal, pou = ref.field("phi"), ref.field("walls_all"), ref.field("pout")
        wal[...] = 0
        pou[...] = 0
        xp = xl - 2
        for x in range(xl):
            height = 80.0 + 25.0 * np.cos(2 * np.pi * (np.arange(yl) - 37.0 - 3 * x) / yl) + rng.uniform(-2, 2, yl)
            phi[x] = np.tanh(height[:, None] - zz[None, :])
            if x == xp:
                neck = 1 + int(np.argmax(height[1:yl - 1]))
        cols = rng.choice(np.setdiff1d(np.arange(1, yl - 1), [neck]), 40, replace=False)
        for y in [neck] + list(cols[:12]):  # a lower blue band: an earlier crossing
            phi[xp, y, 20:24] = -0.5
        for y in cols[12:20]:  # the crossing's upper node is a wall node
            zc = int(np.argmax(phi[xp, y] < 0))
            wal[xp, y, zc] = 10
        for y in cols[20:28]:  # its lower node is a pout node
            zc = int(np.argmax(phi[xp, y] < 0))
            pou[xp, y, zc - 1] = 1
        for y in cols[28:32]:  # no crossing
            phi[xp, y] = 0.5
        h, y0 = hm.hmin(phi.copy(), wal.copy(), pou.copy())
        assert 2 < y0 < yl - 1 and h < zl
        sc[5], sc[6] = tr, float(t)
        ref.call("output3")
        lines = ref.hmin_lines()
        assert lines[-1] == hm.line(t, tr, h, y0, RAD_C, lamda_d), (t, lines[-1])
        if t == 0:
            assert lines[0].strip() == "VARIABLES=t,t*,h,y0" and len(lines) == 2
