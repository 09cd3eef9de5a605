"""GPU: the scatter form of the phase field (periodic FP32/FP16 factorised
boxes).  The step kernel adds each node's post-collision red and blue
populations to the sums of its 19 downstream nodes in 32-bit fixed point, so
the next step needs no moments pass (calmacros' sums, S:792-799, taken from
the writers' side).

* Against the gather form (a moments pass per step, CGL_SCATTER=0): the two
  differ only in the rounding of rho_R, so after the first scattered step
  phi and u agree to 1e-6 (L2-relative).  Over 200 steps both stay within the
  parity tolerances of the FP64 path (rounding differences grow with the
  flow, so the long-run check is against FP64, not between the two forms).
* Fixed-point sums are exact, so results do not depend on the workgroup box
  decomposition: every x chunk, including chunks of one plane and a single
  chunk, gives bit-identical states.  Boxes of one workgroup tile in y or z
  (the sums wrap onto the box itself) agree with the gather form.
* Reading the state mid-run (calmacros, derivatives, download, diagnostics)
  and graph replay do not change the trajectory.
"""
import os

import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import rel_err, small_rp

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _lattice(cfg, precision="f32", scatter=True, graphs=True, **kw):
    saved = {k: os.environ.get(k) for k in ("CGL_SCATTER", "CGL_GRAPHS")}
    os.environ["CGL_SCATTER"] = "1" if scatter else "0"
    os.environ["CGL_GRAPHS"] = "1" if graphs else "0"
    try:
        lat = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                                   rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"],
                                   layout="fs", **kw)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return lat


def _fields(lat):
    m = lat.calmacros()
    return m["phi"], np.stack([m["ux"], m["uy"], m["uz"]])


@pytest.fixture(scope="module")
def rp():
    return small_rp(nx=32, ny=32, nz=128, nu_b=0.1)


def test_scatter_form_selection(rp):
    cfg, _, _ = rp
    assert _lattice(cfg).scatter_form()
    assert _lattice(cfg, "f16").scatter_form()
    assert not _lattice(cfg, scatter=False).scatter_form()
    assert not _lattice(cfg, "f64").scatter_form()
    assert _lattice(dict(cfg, ny=12, nz=96)).scatter_form()  # partial workgroup tiles
    assert not _lattice(dict(cfg, ny=17)).scatter_form()  # a one-row tile: faces on one thread
    assert not _lattice(dict(cfg, nz=129)).scatter_form()  # a one-lane tile


def test_scatter_matches_gather_after_three_steps(rp):
    """Steps 1 and 2 take phi from a moments pass in both forms (layout
    conversion, then the scatter form's first step); step 3 uses the
    scattered sums, which differ from the gathered ones in rounding only."""
    cfg, ri, bi = rp
    out = []
    for sc in (True, False):
        lat = _lattice(cfg, scatter=sc, xchunk=8)
        assert lat.scatter_form() == sc
        lat.upload(ri, bi)
        lat.step(3)
        out.append(_fields(lat))
    e_phi, e_u = rel_err(out[0][0], out[1][0]), rel_err(out[0][1], out[1][1])
    _print("scatter vs gather, 3 steps: phi", e_phi, "u", e_u)
    assert 0 < e_phi[0] < 1e-6 and e_u[0] < 1e-6


@pytest.mark.parametrize("prec,tol_phi,tol_u", [("f32", 1e-5, 1e-5), ("f16", 5e-3, 5e-2)])
def test_scatter_and_gather_against_f64(prec, tol_phi, tol_u):
    """Both forms against the FP64 path after 200 steps on the Rayleigh-Plateau
    box of the parity tests, at the parity tolerances; the scatter form is not
    further from FP64 than the gather form by more than a factor 2."""
    cfg, ri, bi = small_rp()
    ref = _lattice(cfg, "f64")
    ref.upload(ri, bi)
    ref.step(200)
    p64, u64 = _fields(ref)
    errs = {}
    for sc in (True, False):
        lat = _lattice(cfg, prec, scatter=sc)
        lat.upload(ri, bi)
        lat.step(200)
        phi, u = _fields(lat)
        errs[sc] = (rel_err(phi, p64)[0], rel_err(u, u64)[0])
    _print(prec, "vs f64 (phi, u): scatter", errs[True], "gather", errs[False])
    assert errs[True][0] < tol_phi and errs[True][1] < tol_u
    assert errs[True][0] <= 2 * errs[False][0] + 1e-7 and errs[True][1] <= 2 * errs[False][1] + 1e-7


@pytest.mark.parametrize("ny,nz", [(16, 128), (12, 96)])
def test_scatter_independent_of_box_decomposition(ny, nz):
    """x chunks 1, 2, 3 (last chunk one plane), 8, 40 (one chunk) and 64 on a
    40-plane box, with whole and partial y/z tiles: bit-identical states."""
    cfg, ri, bi = small_rp(nx=40, ny=ny, nz=nz, r0=min(ny, 16) / 3.2, nu_b=0.1)
    ref = None
    for xc in (8, 1, 2, 3, 40, 64):
        _print("decomposition xchunk", xc)
        lat = _lattice(cfg, xchunk=xc)
        lat.upload(ri, bi)
        lat.step(9)
        r, b = lat.download()
        if ref is None:
            ref = (r, b)
        else:
            assert np.array_equal(r, ref[0]) and np.array_equal(b, ref[1]), xc


@pytest.mark.parametrize("shape", [(6, 8, 64), (10, 8, 128), (12, 24, 64), (8, 12, 96), (6, 10, 130), (5, 6, 40),
                                   (7, 12, 95)])
def test_scatter_single_tile_axes(shape):
    """One workgroup tile along y (ny = 8) and/or z (nz = 64): the box-leaving
    sums wrap onto the box itself.  Partial tiles: a 4-row and a 32-lane tile
    (12 x 96), a 2-row and a 2-lane tile (10 x 130), one partial tile on both
    axes (6 x 40); an odd nz (95), whose finishing pass takes one node per
    thread (even nz: two)."""
    nx, ny, nz = shape
    cfg, ri, bi = small_rp(nx=nx, ny=ny, nz=nz, r0=2.5, nu_b=0.1)
    out = []
    for sc in (True, False):
        lat = _lattice(cfg, scatter=sc, xchunk=4)
        lat.upload(ri, bi)
        lat.step(5)
        out.append(_fields(lat))
    e_phi, e_u = rel_err(out[0][0], out[1][0]), rel_err(out[0][1], out[1][1])
    _print("single tile", shape, "phi", e_phi, "u", e_u)
    assert e_phi[0] < 1e-6 and e_u[0] < 1e-5


def test_observation_does_not_change_the_run(rp):
    cfg, ri, bi = rp
    a = _lattice(cfg, xchunk=8)
    a.upload(ri, bi)
    a.step(24)
    ra, ba = a.download()
    b = _lattice(cfg, xchunk=8)
    b.upload(ri, bi)
    b.step(7)
    b.calmacros()
    b.step(5)
    b.derivatives()
    b.download()
    b.diagnostics()
    b.step(12)
    rb, bb = b.download()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
    c = _lattice(cfg, xchunk=8, graphs=False)
    c.upload(ri, bi)
    for _ in range(24):
        c.step(1)
    rc, bc = c.download()
    assert np.array_equal(ra, rc) and np.array_equal(ba, bc)


def test_timed_steps_leave_a_readable_state(rp):
    cfg, ri, bi = rp
    a = _lattice(cfg, xchunk=8)
    a.upload(ri, bi)
    a.step(3)
    a.time_steps(4)
    ra, ba = a.download()
    b = _lattice(cfg, xchunk=8)
    b.upload(ri, bi)
    b.step(3 + 8)  # time_steps runs its steps twice (loop timing, then per-kernel timing)
    rb, bb = b.download()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)


def _ball_state(cfg, rho_in):
    """Pure blue at rho0 around a pure red ball (radius 6, centred) of
    density rho_in, at rest: the populations are w_i times the densities."""
    nx, ny, nz = cfg["nx"], cfg["ny"], cfg["nz"]
    x, y, z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ball = (x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2 <= 36.0
    w = ic.W
    return np.where(ball, rho_in, 0.0)[..., None] * w, np.where(ball, 0.0, cfg["rho0"])[..., None] * w


@pytest.mark.parametrize("rho_in", [3.0, 5.0])
def test_density_beyond_the_fixed_point_range_is_reported(rp, rho_in):
    """The scatter form sums colour densities in 32-bit fixed point at 2^29
    units per rho0.  A red ball at 3 rho0 steps normally; one at 5 rho0
    wraps the sums, and sync, calmacros, download and diagnostics report
    the range error until a new state is installed."""
    cfg, ri, bi = rp
    lat = _lattice(cfg)
    assert lat.scatter_form()
    lat.upload(*_ball_state(cfg, rho_in))
    lat.step(4)
    if rho_in < 4.0:
        lat.sync()
        m = lat.calmacros()
        _print("ball at", rho_in, "rho0: max rho after 4 steps", float(m["rho"].max()))
        assert np.isfinite(m["phi"]).all() and float(m["rho"].max()) > 2.5
        return
    for call in (lat.sync, lat.calmacros, lat.download, lat.diagnostics):
        with pytest.raises(symmetry_amd.CglError) as e:
            call()
        assert e.value.code == symmetry_amd.ERR_RANGE, call
    lat.upload(ri, bi)
    lat.step(4)
    lat.sync()
    phi, u = _fields(lat)
    assert np.isfinite(phi).all() and np.isfinite(u).all()


def _shaped(cfg, tile_rows=0, remap=None, xchunk=8):
    saved = os.environ.get("CGL_XCD_REMAP")
    if remap is not None:
        os.environ["CGL_XCD_REMAP"] = str(remap)
    try:
        return _lattice(cfg, xchunk=xchunk, tile_rows=tile_rows)
    finally:
        if saved is None:
            os.environ.pop("CGL_XCD_REMAP", None)
        else:
            os.environ["CGL_XCD_REMAP"] = saved


@pytest.mark.parametrize("shape", [(32, 24, 128), (20, 12, 96)])
def test_launch_shape_does_not_change_the_run(shape):
    """Tile heights 4, 8 and 16 rows and the three XCD tile orders give the
    same state bit for bit (the second shape has partial y and z tiles)."""
    nx, ny, nz = shape
    cfg, ri, bi = small_rp(nx=nx, ny=ny, nz=nz, r0=5.0, amps=(0.5,), nu_b=0.1)
    ref = _shaped(cfg)
    ref.upload(ri, bi)
    ref.step(9)
    ra, ba = ref.download()
    for rows in (4, 8, 16):
        for remap in (0, 1, 2):
            lat = _shaped(cfg, tile_rows=rows, remap=remap)
            assert lat.scatter_form()
            tp = lat.tuned_params()
            assert tp["tile_rows"] == rows and tp["xcd_order"] == remap and not tp["tuning"]
            lat.upload(ri, bi)
            lat.step(9)
            r, b = lat.download()
            _print(shape, "rows", rows, "order", remap, "max |d|", float(np.max(np.abs(r - ra))))
            assert np.array_equal(r, ra) and np.array_equal(b, ba), (rows, remap)
            lat.close()


def test_autotuner_reports_its_choice_and_keeps_the_run(rp):
    """xchunk = 0: the first steps time the x chunks, then the tile orders,
    then the tile heights; afterwards the handle reports its choice, and the
    run equals one with that shape fixed from the start bit for bit."""
    cfg, ri, bi = rp
    lat = _lattice(cfg, xchunk=0)
    lat.upload(ri, bi)
    steps = 0
    while lat.tuned_params()["tuning"]:
        lat.step(2)
        steps += 2
        assert steps <= 60
    tp = lat.tuned_params()
    _print("tuned after", steps, "steps:", tp)
    assert tp["xchunk"] in (2, 4, 8, 16, 32, 64) and tp["xcd_order"] in (0, 1, 2) and tp["tile_rows"] in (4, 8, 16)
    lat.step(4)
    ra, ba = lat.download()
    fixed = _shaped(cfg, tile_rows=tp["tile_rows"], remap=tp["xcd_order"], xchunk=tp["xchunk"])
    fixed.upload(ri, bi)
    fixed.step(steps + 4)
    rb, bb = fixed.download()
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
