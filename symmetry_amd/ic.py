"""Host-side initial conditions and workload definitions.

The liquid-cylinder state follows the reference's ``initialization``
(Source1.f90 S:850-896): tanh profile phi = tanh(6 beta 0.134 (R - r)), a
Laplace density jump 3 sigma / R inside the liquid, colour densities
rho (1 +- phi)/2 and SRT equilibria at rest (``equilf_SRT``, S:833-848).  The
reference puts the cylinder axis along y through the box corner with one
cosine mode; the workloads put it along z through the box centre with one or
several modes.  Arithmetic follows the reference's operation order so that
this generator reproduces the reference's own initial state exactly when
given the reference's geometry.
"""
from __future__ import annotations

import math

import numpy as np

Q = 19
W = np.array([1.0 / 3.0] + [1.0 / 18.0] * 6 + [1.0 / 36.0] * 12)
_tanh = np.frompyfunc(math.tanh, 1, 1)


def cylinder_state(nx, ny, nz, *, axis, center, r0, amps, wavelength, phase0, sigma, rho0, beta=0.7,
                   exact=True):
    """Populations (ri, bi) [x, y, z, i] and phi [x, y, z] of a perturbed
    liquid cylinder.

    axis       'y' (reference) or 'z' (workloads): the cylinder axis.
    center     the two transverse coordinates of the axis, 1-based lattice
               units, in (x, z) order for axis 'y' and (x, y) for axis 'z'.
    amps       a_m of R(s) = r0 + sum_m a_m cos(2 pi m (s - phase0) / wavelength).
    exact      use libm tanh per node (bit-reproducible); False uses numpy's
               vectorised tanh (faster, may differ in the last ulp).
    """
    x = np.arange(1, nx + 1, dtype=np.float64)[:, None, None]
    y = np.arange(1, ny + 1, dtype=np.float64)[None, :, None]
    z = np.arange(1, nz + 1, dtype=np.float64)[None, None, :]
    if axis == "y":
        dis = (x - center[0]) ** 2 + (z - center[1]) ** 2
        s = y
    elif axis == "z":
        dis = (x - center[0]) ** 2 + (y - center[1]) ** 2
        s = z
    else:
        raise ValueError(axis)
    cos = (lambda v: np.vectorize(math.cos, otypes=[np.float64])(v)) if exact else np.cos
    dis_x = np.zeros_like(s)
    for m, a in enumerate(amps, start=1):
        dis_x = dis_x + a * cos(2.0 * math.pi * m * (s - phase0) / wavelength)
    star = r0 + dis_x - np.sqrt(dis)
    star = np.broadcast_to(star, (nx, ny, nz))
    arg = star * 6.0 * beta * 0.134
    phi = _tanh(arg).astype(np.float64) if exact else np.tanh(arg)
    plus = np.broadcast_to(sigma / (r0 + dis_x) * 3.0, (nx, ny, nz))
    rho = np.where(phi > 0.0, rho0 + plus, rho0)
    rho_r = rho * (phi + 1.0) / 2.0
    rho_b = rho - rho_r
    ri = (rho_r[..., None] * W) * 1.0
    bi = (rho_b[..., None] * W) * 1.0
    return ri, bi, phi


# ---------------------------------------------------------------- workloads

def rp_config1():
    """Plain FP64 Rayleigh-Plateau, 32x32x128 periodic, R=8, 10% cosine
    perturbation with the box length as wavelength, sigma=A=0.01, beta=0.7,
    tau=1.0 for both fluids, rho=1, at rest; 2000 steps."""
    return dict(nx=32, ny=32, nz=128, r0=8.0, amps=(0.8,), sigma=0.01, rho0=1.0,
                nu_r=(1.0 - 0.5) / 3.0, nu_b=(1.0 - 0.5) / 3.0, beta=0.7, steps=2000)


def rp_config3_amps(seed=1234, r0=40.0):
    """Amplitudes of the 8 modes of the 256x256x512 workload: uniform in
    [0, 0.02 R] from numpy's PCG64 generator with the given seed."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 0.02 * r0, 8)


def rp_config3():
    """Precision-sweep Rayleigh-Plateau, 256x256x512 periodic, R=40, 8 modes,
    tau_red=1.0, tau_blue=0.7, A=sigma=0.01, beta=0.7; 1000 steps."""
    return dict(nx=256, ny=256, nz=512, r0=40.0, amps=tuple(rp_config3_amps()), sigma=0.01, rho0=1.0,
                nu_r=(1.0 - 0.5) / 3.0, nu_b=(0.7 - 0.5) / 3.0, beta=0.7, steps=1000)


def workload_state(cfg, exact=True, nx=None, ny=None, nz=None):
    """Host initial state of a z-axis cylinder workload (optionally resized)."""
    nx = nx or cfg["nx"]
    ny = ny or cfg["ny"]
    nz = nz or cfg["nz"]
    return cylinder_state(nx, ny, nz, axis="z", center=(nx / 2 + 0.5, ny / 2 + 0.5), r0=cfg["r0"],
                          amps=cfg["amps"], wavelength=float(nz), phase0=1.0, sigma=cfg["sigma"],
                          rho0=cfg["rho0"], beta=cfg["beta"], exact=exact)


def flat_interface_state(nx, ny, nz, *, z0, rho0, beta=0.7, drop=None, sigma=0.0, amp=0.0, exact=False):
    """Red fluid below the surface z = z0 + amp cos(2 pi (x - 1) / nx)
    (1-based lattice units), blue above, phi = tanh(6 beta 0.134 (z0 - z)) at
    rest, optionally with a red drop ``drop = (cx, cy, cz, r)`` (union of the
    two red regions, with the Laplace jump 3 sigma / r inside the drop).  Same
    profile and equilibria as ``cylinder_state``; ``exact`` uses libm tanh
    per node (bit-reproducible)."""
    x = np.arange(1, nx + 1, dtype=np.float64)[:, None, None]
    y = np.arange(1, ny + 1, dtype=np.float64)[None, :, None]
    z = np.arange(1, nz + 1, dtype=np.float64)[None, None, :]
    zs = z0 + amp * np.cos(2.0 * math.pi * (x - 1.0) / nx)
    star = np.broadcast_to(zs - z, (nx, ny, nz)).copy()
    rho = np.full((nx, ny, nz), rho0)
    if drop is not None:
        cx, cy, cz, r = drop
        sd = r - np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2)
        rho = np.where(sd > 0.0, rho0 + 3.0 * sigma / r, rho0)
        star = np.maximum(star, sd)
    arg = star * 6.0 * beta * 0.134
    phi = _tanh(arg).astype(np.float64) if exact else np.tanh(arg)
    rho_r = rho * (phi + 1.0) / 2.0
    rho_b = rho - rho_r
    return (rho_r[..., None] * W) * 1.0, (rho_b[..., None] * W) * 1.0, phi


def sphere_particles(centers, rad, rho_p, cita_deg, sigma):
    """Rigid spheres of the reference's kind (S:76-84): mass 4/3 pi rho_p R^3,
    moment of inertia 2/5 m R^2, contact angle in radians and the repulsion
    scale Fg = -2 pi R sigma."""
    centers = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    n = len(centers)
    rad_a = np.broadcast_to(np.asarray(rad, dtype=np.float64), (n,)).copy()
    rho_a = np.broadcast_to(np.asarray(rho_p, dtype=np.float64), (n,))
    mass = 4.0 / 3.0 * rho_a * math.pi * rad_a ** 3
    inertia = 2.0 / 5.0 * mass * rad_a ** 2
    cita = np.broadcast_to(np.asarray(cita_deg, dtype=np.float64), (n,)) / 180.0 * math.pi
    fg = -math.pi * 2.0 * float(rad_a[0]) * sigma
    return dict(center=centers, rad=rad_a, mass=mass, inertia=inertia, cita=cita.copy(), fg=fg)


def particle_config2(n=96, rad=10.0):
    """Single moving sphere at a flat interface: n^3 box, walls at z = 1 and
    z = n (ghost layers), periodic in x and y, red below z = n/2, sphere of
    radius ``rad`` at the box centre with density 1.5 and a 60 degree contact
    angle, tau = 0.8, A = 0.005, beta = 0.7; 5000 steps."""
    tau = 0.8
    return dict(nx=n, ny=n, nz=n, bc=(0, 0, 0, 0, 1, 1), z0=n / 2.0, center=(n / 2.0, n / 2.0, n / 2.0),
                rad=rad, rho_p=1.5, cita_deg=60.0, sigma=0.005, rho0=1.0, nu_r=(tau - 0.5) / 3.0,
                nu_b=(tau - 0.5) / 3.0, beta=0.7, steps=5000)


def ensemble_config4(seed=7, nbox=64):
    """Batched ensemble of uneven periodic Rayleigh-Plateau boxes.  Per box, in
    this draw order from numpy's PCG64 generator with the given seed:
    nx = ny from {32, 48, 64, 96}; nz from {64, 96, ..., 512}; cylinder radius
    R uniform in [0.15 nx, 0.3 nx]; one cosine mode of wavelength nz with
    amplitude uniform in [0.01, 0.1] R; tau uniform in [0.6, 1.2] for both
    fluids.  sigma = A = 0.01, beta = 0.7, rho0 = 1 as in the other
    Rayleigh-Plateau workloads; 2000 steps."""
    rng = np.random.default_rng(seed)
    boxes = []
    for _ in range(nbox):
        nx = int(rng.choice([32, 48, 64, 96]))
        nz = int(rng.choice(np.arange(64, 513, 32)))
        r0 = float(rng.uniform(0.15 * nx, 0.3 * nx))
        amp = float(rng.uniform(0.01, 0.1)) * r0
        tau = float(rng.uniform(0.6, 1.2))
        nu = (tau - 0.5) / 3.0
        boxes.append(dict(nx=nx, ny=nx, nz=nz, r0=r0, amps=(amp,), tau=tau, nu_r=nu, nu_b=nu, sigma=0.01,
                          rho0=1.0, beta=0.7))
    return boxes


def particle_config5(seed=99, n=512, npar=200, rad=8.0, band=24.0):
    """Dense particle-laden interface: n^3 box, walls at z = 1 and z = n,
    periodic in x and y, red fluid below z = n/2.  npar spheres of radius
    ``rad`` placed without overlap by random sequential addition from numpy's
    PCG64 generator with the given seed: centres uniform in x, y over the box
    and in z within +-band of the interface, a draw rejected when it comes
    closer than 2 rad + 1 to an accepted sphere (minimum image in x and y);
    then densities uniform in [0.8, 2.0] and contact angles uniform in
    [30, 150] degrees, in particle order.  tau = 0.8, A = 0.005, beta = 0.7;
    2000 steps."""
    rng = np.random.default_rng(seed)
    z0 = n / 2.0
    centers = []
    while len(centers) < npar:
        c = np.array([rng.uniform(1.0, n + 1.0), rng.uniform(1.0, n + 1.0), rng.uniform(z0 - band, z0 + band)])
        ok = True
        for d in centers:
            dx = np.abs(c - d)
            dx[:2] = np.minimum(dx[:2], n - dx[:2])
            if np.sqrt((dx * dx).sum()) < 2.0 * rad + 1.0:
                ok = False
                break
        if ok:
            centers.append(c)
    rho_p = rng.uniform(0.8, 2.0, npar)
    cita = rng.uniform(30.0, 150.0, npar)
    tau = 0.8
    return dict(nx=n, ny=n, nz=n, bc=(0, 0, 0, 0, 1, 1), z0=z0, centers=np.array(centers), rad=rad, rho_p=rho_p,
                cita_deg=cita, sigma=0.005, rho0=1.0, nu_r=(tau - 0.5) / 3.0, nu_b=(tau - 0.5) / 3.0, beta=0.7,
                steps=2000)
