"""The Fortran driver (symmetry_amd/cgl_driver, built from fortran/) calls
libcgl through its bind(C) module exactly where the reference's time loop
calls its subroutines.  Run it on a small box and check its dumped fields
bit for bit against the FP64 oracle fed the same initial state, and its
Tecplot dump against the reference's output_results format (S:910-921)."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import energy as en
from oracle import hmin as hm
from symmetry_amd import ic
from tests.helpers import oracle_from

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "symmetry_amd", "cgl_driver")


def test_fortran_driver_matches_oracle(tmp_path):
    nx, ny, nz, steps = 16, 12, 48, 3
    nml = tmp_path / "in.nml"
    nml.write_text(f"&cgl\n xl={nx}, yl={ny}, zl={nz}, nsteps={steps}, out_every={steps}, prec=0,\n"
                   f" sigma=0.01d0, rho0=1.0d0, tau_R=1.0d0, tau_B=0.8d0, beta=0.7d0, Rad_c=4.0d0,\n"
                   f" nmodes=1, amp(1)=0.4d0, dump_binary=.true., write_tecplot=.true., outdir='{tmp_path}'\n/\n")
    print("driver", nx, ny, nz, steps, flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    raw = np.fromfile(tmp_path / f"FIELDS_{steps}.bin", dtype=np.float64).reshape(7, nx, ny, nz)
    cfg = dict(nx=nx, ny=ny, nz=nz, r0=4.0, amps=(0.4,), sigma=0.01, rho0=1.0, nu_r=(1.0 - 0.5) / 3.0, nu_b=(0.8 - 0.5) / 3.0,
               beta=0.7)
    ri, bi, _ = ic.workload_state(cfg, exact=True)
    st = oracle_from(cfg, ri, bi)
    st.step(steps)
    st.call("calmacros")
    for k, name in enumerate(("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")):
        assert np.array_equal(raw[k], st.arrays[name]), name
    lines = (tmp_path / f"OUT_{steps}.dat").read_text().splitlines()
    assert "VARIABLES=X,Y,Z,WALL,RHO_N,U,V,W,RHO" in lines[1]
    first = lines[3].split()
    assert [int(v) for v in first[:4]] == [1, 1, 1, 0]
    assert abs(float(first[4]) - st.phi[0, 0, 0]) < 1e-8
    assert len(lines) == 3 + nx * ny * nz
    # the energy budget (output4, S:1719-1797) of a box taller than wide and
    # without particles: the reads at x index z stop at z = xl
    e = en.energy(raw[6], raw[0], raw[3], raw[4], raw[5], np.zeros((nx, ny, nz), np.int32), rho0=1.0, sigma=0.01,
                  rad_c=4.0, up=[], rp=[], tension=[], tension_t=[], mass=[], inertia=[], center0=[], shb=[],
                  wcap_old=0.0)
    elines = (tmp_path / "Energy.dat").read_text().splitlines()
    tr = (4.0 ** 3 * 1.0 / 0.01) ** 0.5
    assert len(elines) == 3 and elines[2] == en.line(steps, tr, e) and e["Esur"] > 0


def test_fortran_driver_particle_at_interface(tmp_path):
    """The particle workload's set-up (z walls, flat interface, one sphere of
    density 1.5 and a 60 degree contact angle) through the Fortran driver
    equals the same run through the ctypes binding, bit for bit: the driver's
    initial state, particle parameters (S:81-84) and calls are the ones the
    library tests check against the oracle."""
    import symmetry_amd
    n, steps, rad = 32, 8, 5.0
    c = (16.3, 15.8, 16.4)
    nml = tmp_path / "in.nml"
    nml.write_text(f"&cgl\n xl={n}, yl={n}, zl={n}, nsteps={steps}, out_every={steps}, prec=0,\n"
                   f" sigma=0.005d0, rho0=1.0d0, tau_R=0.8d0, tau_B=0.8d0, beta=0.7d0,\n"
                   f" ic='interface', z0=16.0d0, zwalls=.true., npar=1, rad_p={rad}d0, rho_p=1.5d0, cita_deg=60.0d0,\n"
                   f" dump_binary=.true., write_tecplot=.true., outdir='{tmp_path}'\n/\n"
                   f"&particles\n px={c[0]}d0, py={c[1]}d0, pz={c[2]}d0\n/\n")
    print("driver particle", n, steps, flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    raw = np.fromfile(tmp_path / f"FIELDS_{steps}.bin", dtype=np.float64).reshape(7, n, n, n)
    rows = [l.split() for l in (tmp_path / "PARTICLES.dat").read_text().splitlines()[1:]]
    last = [r_ for r_ in rows if int(r_[0]) == steps][0]
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=16.0, rho0=1.0, beta=0.7, exact=True)
    parts = ic.sphere_particles([c], rad, 1.5, 60.0, 0.005)
    lat = symmetry_amd.Lattice(n, n, n, precision="f64", sigma=0.005, rho0=1.0, nu_r=(0.8 - 0.5) / 3.0,
                               nu_b=(0.8 - 0.5) / 3.0, beta=0.7, bc=(0, 0, 0, 0, 1, 1), npar=1)
    lat.upload(ri, bi)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"])
    lat.step(steps)
    mac = lat.calmacros()
    for k, name in enumerate(("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")):
        assert np.array_equal(raw[k], mac[name]), name
    got = lat.get_particles()
    vals = [float(v) for v in last[2:]]
    assert vals[0:3] == list(got["center"][0]) and vals[3:6] == list(got["up"][0])
    assert vals[6:9] == list(got["force"][0]) and vals[9:12] == list(got["fcap"][0])
    masks = lat.masks()
    lines = (tmp_path / f"OUT_{steps}.dat").read_text().splitlines()
    wall_col = np.array([int(l.split()[3]) for l in lines[3:]]).reshape(n, n, n)  # (z, y, x), x fastest
    assert np.array_equal(wall_col.transpose(2, 1, 0), masks["walls_all"])


def test_fortran_driver_reference_faces_with_particles(tmp_path):
    """ref_faces: the reference's own box (walls at x = 1 and z = 1, mirrors at
    x = xl, y = 1, y = yl, z = zl) with particles on the mirror planes, its
    repulsion planes, Lagrange box and symmetry-plane rules; the Fortran
    driver equals the same run through the ctypes binding bit for bit, and
    its Energy.dat lines (output4) equal the oracle's restatement of output4
    on the binding's state."""
    import symmetry_amd
    n, steps, rad = 32, 6, 5.0
    cs = [(n - 0.5, 15.8, 16.4), (14.3, 1.5, 17.2)]
    nml = tmp_path / "in.nml"
    nml.write_text(f"&cgl\n xl={n}, yl={n}, zl={n}, nsteps={steps}, out_every={steps}, prec=0,\n"
                   f" sigma=0.005d0, rho0=1.0d0, tau_R=0.8d0, tau_B=0.8d0, beta=0.7d0,\n"
                   f" ic='interface', z0=16.0d0, ref_faces=.true., npar=2,\n"
                   f" rad_p={rad}d0, rho_p=1.5d0,\n"
                   f" cita_deg=90.0d0, dump_binary=.true., write_tecplot=.false., outdir='{tmp_path}'\n/\n"
                   f"&particles\n px={cs[0][0]}d0, {cs[1][0]}d0, py={cs[0][1]}d0, {cs[1][1]}d0,\n"
                   f" pz={cs[0][2]}d0, {cs[1][2]}d0\n/\n")
    print("driver reference faces", n, steps, flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    raw = np.fromfile(tmp_path / f"FIELDS_{steps}.bin", dtype=np.float64).reshape(7, n, n, n)
    ri, bi, _ = ic.flat_interface_state(n, n, n, z0=16.0, rho0=1.0, beta=0.7, exact=True)
    parts = ic.sphere_particles(cs, rad, 1.5, 90.0, 0.005)
    W, M = symmetry_amd.BC_WALL, symmetry_amd.BC_MIRROR
    lat = symmetry_amd.Lattice(n, n, n, precision="f64", sigma=0.005, rho0=1.0, nu_r=(0.8 - 0.5) / 3.0,
                               nu_b=(0.8 - 0.5) / 3.0, beta=0.7, bc=(W, M, M, M, W, M), npar=2)
    lat.upload(ri, bi)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"])
    lat.set_particle_geometry((1.5, n - 1.5, 1.5, n - 1.5, 1.5, n - 0.5), (1,) * 6, (0.0, 1.5, 0.0),
                              (n - 0.5, n - 0.5, n - 0.5), (0, 0, 1, 1, 0, 1, 0, 0, 1), 1)

    def energy(wcap_old):
        """output4 restated from the binding's state (which equals the
        driver's bit for bit); the driver's Rad_c is its default 8."""
        m, mk, g = lat.calmacros(), lat.masks(), lat.get_particles()
        return en.energy(m["phi"], m["rho"], m["ux"], m["uy"], m["uz"], mk["walls_all"], rho0=1.0, sigma=0.005,
                         rad_c=8.0, up=g["up"], rp=g["rp"], tension=g["fcap"], tension_t=g["tcap"],
                         mass=parts["mass"], inertia=parts["inertia"], center0=np.array(cs), shb=g["shb"],
                         wcap_old=wcap_old)

    tr = math.sqrt(8.0 ** 3 * 1.0 / 0.005)

    def hmin_line(t):
        """output3 restated from the binding's state (Rad_c 8, lamda_d = (yl - 2) 2)."""
        m, mk = lat.calmacros(), lat.masks()
        h, y0 = hm.hmin(m["phi"], mk["walls_all"], mk["pout"])
        return hm.line(t, tr, h, y0, 8.0, (n - 2) * 2.0)

    e0, h0 = energy(0.0), hmin_line(0)
    lat.step(steps)
    e1, h1 = energy(e0["wcap_old"]), hmin_line(steps)
    mac = lat.calmacros()
    for k, name in enumerate(("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")):
        assert np.array_equal(raw[k], mac[name]), name
    got = lat.get_particles()
    rows = [l.split() for l in (tmp_path / "PARTICLES.dat").read_text().splitlines()[1:]]
    last = [r_ for r_ in rows if int(r_[0]) == steps]
    for p, row in enumerate(last):
        vals = [float(v) for v in row[2:]]
        assert vals[0:3] == list(got["center"][p]) and vals[6:9] == list(got["force"][p])
    # the reference's per-particle histories (S:1420-1441) and neck file (S:1673-1715)
    for p in range(2):
        lines = (tmp_path / f"P_{p + 1}.dat").read_text().splitlines()
        assert lines[0].split()[0].startswith("VARIABLES=t,t*,cx") and len(lines) == 3  # t = 0 and t = steps
        cols = [float(v) for v in lines[-1].split()]
        assert len(cols) == 26 and int(cols[0]) == steps
        assert np.allclose(cols[2:5], got["center"][p], atol=1e-6) and np.allclose(cols[11:14], got["force"][p], atol=1e-6)
    hmin = (tmp_path / "hmin.dat").read_text().splitlines()
    assert hmin[0].strip() == "VARIABLES=t,t*,h,y0" and len(hmin) == 3
    assert hmin[1] == h0 and hmin[2] == h1, (hmin, h0, h1)
    # the energy budget (S:1719-1797), both particles on mirror planes (half counts)
    lines = (tmp_path / "Energy.dat").read_text().splitlines()
    assert lines[0].strip() == "VARIABLES=t,t*,Esur,Shb,Efree,Emov,Wcap,Emov_p,Esum" and len(lines) == 3
    assert lines[1] == en.line(0, tr, e0) and lines[2] == en.line(steps, tr, e1)
    assert e1["Shb"] > 0 and e1["Emov_p"] > 0 and e1["Esur"] > 0


def test_fortran_driver_checkpoint_restart(tmp_path):
    """ckpt_every writes CKPT_<t>.bin; a second driver run restarted from
    CKPT_4.bin (restart_step = 4) ends on the same fields and particle state
    bit for bit."""
    n, steps = 24, 10
    common = (f" xl={n}, yl={n}, zl={n}, nsteps={steps}, out_every=5, prec=1,\n"
              f" sigma=0.005d0, rho0=1.0d0, tau_R=0.8d0, tau_B=0.8d0, beta=0.7d0,\n"
              f" ic='interface', z0=12.0d0, zwalls=.true., npar=1,\n"
              f" rad_p=4.0d0, rho_p=1.5d0, cita_deg=90.0d0, dump_binary=.true., write_tecplot=.false.,\n")
    runs = {}
    for tag, extra in (("a", " ckpt_every=4,"), ("b", f" restart='{tmp_path}/a/CKPT_4.bin', restart_step=4,")):
        d = tmp_path / tag
        d.mkdir()
        nml = d / "in.nml"
        nml.write_text("&cgl\n" + common + extra + f" outdir='{d}'\n/\n&particles\n px=12.3d0, py=11.8d0, pz=12.4d0\n/\n")
        print("driver checkpoint run", tag, flush=True)
        r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=120)
        print(r.stdout[-1500:], r.stderr[-1500:])
        assert r.returncode == 0
        runs[tag] = d
    assert (runs["a"] / "CKPT_4.bin").exists() and (runs["a"] / "CKPT_8.bin").exists()
    fa = np.fromfile(runs["a"] / f"FIELDS_{steps}.bin", dtype=np.float64)
    fb = np.fromfile(runs["b"] / f"FIELDS_{steps}.bin", dtype=np.float64)
    assert np.array_equal(fa, fb)
    last = lambda d: [l for l in (d / "PARTICLES.dat").read_text().splitlines()[1:] if int(l.split()[0]) == steps]
    assert last(runs["a"]) == last(runs["b"])


def test_fortran_driver_ensemble_batch(tmp_path):
    """The 64-box ensemble (seed 7) through the driver's batch mode (box_file,
    FP32 factorised): the first two boxes' dumped fields equal the binding's
    batch bit for bit, and every box's BOXES.dat line has its largest speed
    exactly and its masses to summation order."""
    import symmetry_amd
    boxes = ic.ensemble_config4()
    steps = 4
    bf = tmp_path / "boxes.txt"
    bf.write_text("".join(f"{b['nx']} {b['ny']} {b['nz']} {b['tau']!r} {b['tau']!r} {b['r0']!r} {len(b['amps'])} "
                          + " ".join(repr(a) for a in b["amps"]) + "\n" for b in boxes))
    nml = tmp_path / "in.nml"
    nml.write_text(f"&cgl\n nbox={len(boxes)}, box_file='{bf}', nsteps={steps}, out_every={steps}, prec=1, layout=1,\n"
                   f" xchunk=8, sigma=0.01d0, rho0=1.0d0, beta=0.7d0, dump_binary=.true., dump_boxes=2,\n"
                   f" outdir='{tmp_path}'\n/\n")
    print("driver batch, 64 boxes", flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    bt = symmetry_amd.Batch([(b["nx"], b["ny"], b["nz"], b["nu_r"], b["nu_b"]) for b in boxes], precision="f32",
                            layout="fs", sigma=0.01, rho0=1.0, beta=0.7, xchunk=8)
    for k, b in enumerate(boxes):
        bt.init_cylinder(k, b["nx"] / 2 + 0.5, b["ny"] / 2 + 0.5, b["r0"], 6.0 * 0.7 * 0.134, b["amps"])
    bt.step(steps)
    rows = [l.split() for l in (tmp_path / "BOXES.dat").read_text().splitlines()[1:]]
    last = {int(row[1]): [float(v) for v in row[2:]] for row in rows if int(row[0]) == steps}
    assert sorted(last) == list(range(1, len(boxes) + 1))
    names = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")
    for k, b in enumerate(boxes):
        m = bt.calmacros(k)
        umax = float(np.sqrt(m["ux"] ** 2 + m["uy"] ** 2 + m["uz"] ** 2).max())
        mr, mb, um = last[k + 1]
        assert um == umax, (k, um, umax)
        assert abs(mr / m["rho_r"].sum() - 1.0) < 1e-12 and abs(mb / m["rho_b"].sum() - 1.0) < 1e-12, k
        if k < 2:
            raw = np.fromfile(tmp_path / f"BOX_{k + 1}_FIELDS_{steps}.bin", dtype=np.float64)
            raw = raw.reshape(7, b["nx"], b["ny"], b["nz"])
            for j, name in enumerate(names):
                assert np.array_equal(raw[j], m[name]), (k, name)


def test_fortran_driver_dense_particles(tmp_path):
    """The 200-sphere dense workload (512^3, seed 99: per-sphere densities
    and contact angles) through the driver, spheres from particle_file and
    the interface built on the device: its PARTICLES.dat rows equal the
    binding's particle state bit for bit after 0 and 2 steps."""
    import symmetry_amd
    c5 = ic.particle_config5()
    parts = ic.sphere_particles(c5["centers"], c5["rad"], c5["rho_p"], c5["cita_deg"], c5["sigma"])
    n, steps, npar = c5["nx"], 2, len(c5["centers"])
    pf = tmp_path / "spheres.txt"
    pf.write_text("".join(" ".join(repr(float(v)) for v in (*c, c5["rad"], rp, th)) + "\n"
                          for c, rp, th in zip(c5["centers"], c5["rho_p"], c5["cita_deg"])))
    tau = 3.0 * c5["nu_r"] + 0.5
    nml = tmp_path / "in.nml"
    nml.write_text(f"&cgl\n xl={n}, yl={n}, zl={n}, nsteps={steps}, out_every={steps}, prec=1, xchunk=8,\n"
                   f" sigma={c5['sigma']!r}d0, rho0=1.0d0, tau_R={tau!r}d0, tau_B={tau!r}d0, beta=0.7d0,\n"
                   f" ic='interface', ic_on_device=.true., z0={c5['z0']!r}d0, zwalls=.true., npar={npar},\n"
                   f" particle_file='{pf}', write_tecplot=.false., outdir='{tmp_path}'\n/\n")
    print("driver 512^3, 200 spheres", flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    lat = symmetry_amd.Lattice(n, n, n, precision="f32", sigma=c5["sigma"], rho0=1.0, nu_r=c5["nu_r"],
                               nu_b=c5["nu_b"], beta=0.7, bc=c5["bc"], npar=npar, xchunk=8)
    lat.init_interface(c5["z0"], 6.0 * 0.7 * 0.134)
    lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"], parts["fg"])
    rows = [l.split() for l in (tmp_path / "PARTICLES.dat").read_text().splitlines()[1:]]
    for t in (0, steps):
        if t:
            lat.step(t)
        got = lat.get_particles()
        mine = [row for row in rows if int(row[0]) == t]
        assert len(mine) == npar
        for p, row in enumerate(mine):
            vals = [float(v) for v in row[2:]]
            assert vals[0:3] == list(got["center"][p]) and vals[3:6] == list(got["up"][p]), (t, p)
            assert vals[6:9] == list(got["force"][p]) and vals[9:12] == list(got["fcap"][p]), (t, p)
    lat.close()


def test_fortran_driver_neck_and_breakup(tmp_path):
    """The 32x32x128 jet (R = 8, 10 % mode, tau = 1, FP32 factorised) through
    the driver to its breakup: every NECK.dat line (neck radius from the
    device slice sums, its z, the amplitude, phi on the axis at the neck)
    equals the ctypes binding's same run bit for bit, and the driver reports
    the breakup at the same output step."""
    import symmetry_amd
    every, last = 100, 20000
    nml = tmp_path / "in.nml"
    nml.write_text(f"&cgl\n xl=32, yl=32, zl=128, nsteps={last}, out_every={every}, prec=1, layout=1,\n"
                   f" sigma=0.01d0, rho0=1.0d0, tau_R=1.0d0, tau_B=1.0d0, beta=0.7d0, ic_on_device=.true.,\n"
                   f" write_tecplot=.false., outdir='{tmp_path}'\n/\n")
    print("driver jet to breakup", flush=True)
    r = subprocess.run([DRIVER, str(nml)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-600:], r.stderr[-600:])
    assert r.returncode == 0
    rows = [l.split() for l in (tmp_path / "NECK.dat").read_text().splitlines()[1:]]
    assert len(rows) == last // every + 1
    cfg = ic.rp_config1()
    lat = symmetry_amd.Lattice(32, 32, 128, precision="f32", layout="fs", sigma=0.01, rho0=1.0, nu_r=cfg["nu_r"],
                               nu_b=cfg["nu_b"], beta=0.7)
    lat.init_cylinder(16.5, 16.5, 8.0, 6.0 * 0.7 * 0.134, (0.8,))
    breakup = None
    for k, row in enumerate(rows):
        if k:
            lat.step(every)
        rz = np.sqrt(lat.diagnostics(slices=True)["slice_red"] / np.pi)
        phi = lat.calmacros()["phi"]
        ax = (((phi[15, 15] + phi[15, 16]) + phi[16, 15]) + phi[16, 16]) / 4.0  # [x, y] = (16, 16), (16, 17), ... 1-based
        zn = int(np.argmin(rz))
        t = int(row[0])
        assert t == k * every
        assert float(row[2]) == rz[zn] and float(row[3]) == rz.max() and int(row[4]) == zn + 1, (t, row)
        assert float(row[5]) == (rz.max() - rz[zn]) / 2.0 and float(row[6]) == ax.min(), (t, row)
        if breakup is None and ax.min() < 0.0:
            breakup = t
    print("breakup at", breakup, flush=True)
    assert breakup is not None and f" breakup at step {breakup}" in r.stdout
