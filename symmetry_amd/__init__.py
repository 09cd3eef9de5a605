"""symmetry_amd -- MI355X-native per-timestep lattice update of the D3Q19
two-component colour-gradient MRT solver of X-Sheen/Symmetry (Source1.f90).

The product is ``libcgl.so`` (HIP kernels for gfx950 behind the C ABI of
``include/cgl.h``).  This module is a thin ctypes binding of that ABI for
Python callers (tests, benchmarks); the Fortran driver binds the same ABI
through ``fortran/cgl_mod.f90``.  There is no CPU fallback: every call goes to
the GPU library, and importing fails loudly when it has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcgl.so")

BC_PERIODIC, BC_WALL, BC_MIRROR = 0, 1, 2
PREC = {"f64": 0, "f32": 1, "f16": 2}
LAYOUT = {"rb": 0, "fs": 1}
Q = 19


class CglError(RuntimeError):
    """A libcgl call returned an error; ``code`` is its CGL_ERR_* value."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


ERR_ARG, ERR_HIP, ERR_UNSUPPORTED, ERR_RANGE = -1, -2, -3, -4


class Params(ctypes.Structure):
    """Mirror of ``cgl_params`` (include/cgl.h)."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("bc", ctypes.c_int32 * 6),
        ("precision", ctypes.c_int32), ("npar", ctypes.c_int32),
        ("xchunk", ctypes.c_int32), ("device", ctypes.c_int32),
        ("layout", ctypes.c_int32), ("tile_rows", ctypes.c_int32),
        ("sigma", ctypes.c_double), ("rho0", ctypes.c_double),
        ("nu_r", ctypes.c_double), ("nu_b", ctypes.c_double), ("beta", ctypes.c_double),
    ]


class Box(ctypes.Structure):
    """cgl_box of include/cgl.h."""
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("reserved", ctypes.c_int32), ("nu_r", ctypes.c_double), ("nu_b", ctypes.c_double)]


_lib = None

# Every symbol include/cgl.h declares.
EXPORTS = (
    "cgl_create", "cgl_destroy", "cgl_upload_state", "cgl_download_state", "cgl_init_cylinder",
    "cgl_step", "cgl_calmacros", "cgl_derivatives", "cgl_sync", "cgl_time_steps",
    "cgl_device_name", "cgl_last_error", "cgl_probe_copy_bandwidth", "cgl_tuned_xchunk",
    "cgl_set_particles", "cgl_get_particles", "cgl_get_particle_shb", "cgl_download_masks",
    "cgl_batch_create", "cgl_box_upload_state", "cgl_box_init_cylinder", "cgl_box_download_state",
    "cgl_box_calmacros", "cgl_diagnostics", "cgl_init_interface", "cgl_phase_field_form",
    "cgl_set_particle_geometry", "cgl_checkpoint_save", "cgl_checkpoint_load", "cgl_diagnostics_noalloc",
    "cgl_capillary", "cgl_momentum", "cgl_repulsive", "cgl_move", "cgl_set_walls", "cgl_refill",
    "cgl_collide_stream", "cgl_particle_band", "cgl_tuned_params",
)

# The stages of a staged step after cgl_calmacros, in order.
STAGES = ("capillary", "momentum", "repulsive", "move", "set_walls", "refill", "collide_stream")


def lib():
    """Load libcgl.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C {HERE}` "
                              "or __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.cgl_last_error.restype = ctypes.c_char_p
        vp = ctypes.c_void_p
        L.cgl_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params)]
        L.cgl_destroy.argtypes = [vp]
        L.cgl_upload_state.argtypes = [vp, vp, vp]
        L.cgl_download_state.argtypes = [vp, vp, vp]
        L.cgl_init_cylinder.argtypes = [vp] + [ctypes.c_double] * 4 + [ctypes.c_int32, vp]
        L.cgl_step.argtypes = [vp, ctypes.c_int32]
        L.cgl_calmacros.argtypes = [vp] + [vp] * 7
        L.cgl_derivatives.argtypes = [vp] + [vp] * 7
        L.cgl_sync.argtypes = [vp]
        L.cgl_time_steps.argtypes = [vp, ctypes.c_int32] + [ctypes.POINTER(ctypes.c_double)] * 3
        L.cgl_device_name.argtypes = [vp, ctypes.c_char_p, ctypes.c_int32]
        L.cgl_set_particles.argtypes = [vp] + [vp] * 7 + [ctypes.c_double]
        L.cgl_get_particles.argtypes = [vp] + [vp] * 8
        L.cgl_get_particle_shb.argtypes = [vp, vp]
        L.cgl_download_masks.argtypes = [vp] + [vp] * 4
        L.cgl_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Params), ctypes.c_int32, ctypes.POINTER(Box)]
        L.cgl_box_upload_state.argtypes = [vp, ctypes.c_int32, vp, vp]
        L.cgl_box_download_state.argtypes = [vp, ctypes.c_int32, vp, vp]
        L.cgl_box_init_cylinder.argtypes = [vp, ctypes.c_int32] + [ctypes.c_double] * 4 + [ctypes.c_int32, vp]
        L.cgl_box_calmacros.argtypes = [vp, ctypes.c_int32] + [vp] * 7
        L.cgl_diagnostics.argtypes = [vp, vp, vp]
        L.cgl_diagnostics_noalloc.argtypes = [vp, vp, vp]
        L.cgl_init_interface.argtypes = [vp, ctypes.c_double, ctypes.c_double]
        L.cgl_tuned_xchunk.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.cgl_tuned_params.argtypes = [vp] + [ctypes.POINTER(ctypes.c_int32)] * 4
        L.cgl_phase_field_form.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.cgl_set_particle_geometry.argtypes = [vp] + [vp] * 5 + [ctypes.c_int32]
        for st in STAGES:
            getattr(L, "cgl_" + st).argtypes = [vp]
        L.cgl_particle_band.argtypes = [vp] + [vp] * 6
        L.cgl_checkpoint_save.argtypes = [vp, ctypes.c_char_p]
        L.cgl_checkpoint_load.argtypes = [vp, ctypes.c_char_p]
        L.cgl_probe_copy_bandwidth.argtypes = [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                               ctypes.POINTER(ctypes.c_double)]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise CglError(f"libcgl error {rc}: {lib().cgl_last_error().decode()}", rc)


def _ptr(a):
    return None if a is None else a.ctypes.data


def copy_bandwidth(nbytes=2 << 30, reps=20, device=0):
    """Achievable read+write HBM bandwidth (GB/s) of a streaming copy."""
    v = ctypes.c_double()
    _check(lib().cgl_probe_copy_bandwidth(device, int(nbytes), int(reps), ctypes.byref(v)))
    return v.value


class Lattice:
    """One box on one GPU.  Host arrays use the reference's layout with
    0-based indices: populations ``[x, y, z, i]``, scalars ``[x, y, z]``."""

    def __init__(self, nx, ny, nz, precision="f64", sigma=0.01, rho0=1.0, nu_r=1.0 / 6.0, nu_b=1.0 / 6.0,
                 beta=0.7, bc=(BC_PERIODIC,) * 6, xchunk=0, device=0, layout="rb", npar=0, tile_rows=0):
        p = Params()
        p.nx, p.ny, p.nz = nx, ny, nz
        for k, v in enumerate(bc):
            p.bc[k] = v
        p.precision = PREC[precision] if isinstance(precision, str) else int(precision)
        p.npar = npar
        p.xchunk = xchunk
        p.device = device
        p.layout = LAYOUT[layout] if isinstance(layout, str) else int(layout)
        p.tile_rows = tile_rows
        p.sigma, p.rho0, p.nu_r, p.nu_b, p.beta = sigma, rho0, nu_r, nu_b, beta
        self.params = p
        self.shape = (nx, ny, nz)
        self.h = ctypes.c_void_p()
        _check(lib().cgl_create(ctypes.byref(self.h), ctypes.byref(p)))

    def close(self):
        if self.h:
            lib().cgl_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, ri, bi):
        ri = np.ascontiguousarray(ri, dtype=np.float64)
        bi = np.ascontiguousarray(bi, dtype=np.float64)
        assert ri.shape == self.shape + (Q,) and bi.shape == ri.shape
        _check(lib().cgl_upload_state(self.h, _ptr(ri), _ptr(bi)))

    def download(self):
        ri = np.empty(self.shape + (Q,))
        bi = np.empty(self.shape + (Q,))
        _check(lib().cgl_download_state(self.h, _ptr(ri), _ptr(bi)))
        return ri, bi

    def init_cylinder(self, xc, yc, r0, kprof, amps=()):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        _check(lib().cgl_init_cylinder(self.h, xc, yc, r0, kprof, len(a), _ptr(a) if len(a) else None))

    def step(self, n=1):
        _check(lib().cgl_step(self.h, int(n)))

    def sync(self):
        _check(lib().cgl_sync(self.h))

    def calmacros(self):
        names = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")
        out = {k: np.empty(self.shape) for k in names}
        _check(lib().cgl_calmacros(self.h, *(_ptr(out[k]) for k in names)))
        return out

    def derivatives(self):
        names = ("gx", "gy", "gz", "nxv", "nyv", "nzv")
        out = {k: np.empty(self.shape) for k in names}
        out["flag"] = np.empty(self.shape, dtype=np.int32)
        _check(lib().cgl_derivatives(self.h, *(_ptr(out[k]) for k in names), _ptr(out["flag"])))
        return out

    def time_steps(self, n):
        t = [ctypes.c_double() for _ in range(3)]
        _check(lib().cgl_time_steps(self.h, int(n), *(ctypes.byref(v) for v in t)))
        return tuple(v.value for v in t)

    def init_interface(self, z0, kprof):
        _check(lib().cgl_init_interface(self.h, float(z0), float(kprof)))

    def diagnostics(self, slices=False, noalloc=False):
        """FP64 device sums {mass, mass_r, mass_b, jx, jy, jz, ekin} over the
        live fluid nodes, and with slices=True the red area of every z slice.
        noalloc=True takes the allocation-free entry point."""
        sums = np.empty(7)
        area = np.empty(self.shape[2]) if slices else None
        fn = lib().cgl_diagnostics_noalloc if noalloc else lib().cgl_diagnostics
        _check(fn(self.h, _ptr(sums), _ptr(area)))
        out = dict(zip(("mass", "mass_r", "mass_b", "jx", "jy", "jz", "ekin"), sums))
        if slices:
            out["slice_red"] = area
        return out

    def set_particles(self, center, rad, mass, inertia, cita, fg, up=None, rp=None):
        """center, up, rp: [npar, 3] arrays (passed to the library in the
        reference's component-major (npar, 3) Fortran layout)."""
        f = lambda a: None if a is None else np.ascontiguousarray(a, dtype=np.float64)
        t = lambda a: None if a is None else np.ascontiguousarray(np.asarray(a, dtype=np.float64).T)
        args = [t(center), t(up), t(rp), f(rad), f(mass), f(inertia), f(cita)]
        self._keep = args
        _check(lib().cgl_set_particles(self.h, *(_ptr(a) for a in args), float(fg)))

    def set_particle_geometry(self, rep_plane, rep_on, dom_lo=(0.0, 0.0, 0.0), dom_hi=(0.0, 0.0, 0.0),
                              dom_on=(0,) * 9, sym_rules=0):
        """Repulsion planes (6 faces), Lagrange in-domain box (dom_on = low
        on, low strict, high on per axis) and the symmetry-plane rules of the
        reference's own box; call after set_particles."""
        d = lambda a, n: np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(n))
        i = lambda a, n: np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(n))
        args = [d(rep_plane, 6), i(rep_on, 6), d(dom_lo, 3), d(dom_hi, 3), i(dom_on, 9)]
        _check(lib().cgl_set_particle_geometry(self.h, *(_ptr(a) for a in args), int(sym_rules)))

    def get_particles(self):
        n = self.params.npar
        names = ("center", "up", "rp", "force", "torque", "fcap", "tcap", "frep")
        out = {k: np.empty((3, n)) for k in names}
        _check(lib().cgl_get_particles(self.h, *(_ptr(out[k]) for k in names)))
        res = {k: np.ascontiguousarray(v.T) for k, v in out.items()}
        shb = np.empty(n)
        _check(lib().cgl_get_particle_shb(self.h, _ptr(shb)))
        res["shb"] = shb
        return res

    def masks(self):
        names = ("walls", "walls_all", "pin", "pout")
        out = {k: np.empty(self.shape, dtype=np.int32) for k in names}
        _check(lib().cgl_download_masks(self.h, *(_ptr(out[k]) for k in names)))
        return out

    def stage(self, name):
        """Run one stage of a staged step (``STAGES``; calmacros opens it)."""
        if name not in STAGES:
            raise ValueError(name)
        _check(getattr(lib(), "cgl_" + name)(self.h))

    def particle_band(self):
        """The FP64 band fields {rho, rho_r, rho_b, ux, uy, uz} the particle
        stages read and write (defined within R + 3 of the centres)."""
        names = ("rho", "rho_r", "rho_b", "ux", "uy", "uz")
        out = {k: np.empty(self.shape) for k in names}
        _check(lib().cgl_particle_band(self.h, *(_ptr(out[k]) for k in names)))
        return out

    def tuned_xchunk(self):
        """(x chunk in use, True while the autotuner is still running)."""
        xc, tu = ctypes.c_int32(), ctypes.c_int32()
        _check(lib().cgl_tuned_xchunk(self.h, ctypes.byref(xc), ctypes.byref(tu)))
        return xc.value, bool(tu.value)

    def tuned_params(self):
        """{xchunk, xcd_order, tile_rows, tuning}: the launch shape in use."""
        v = [ctypes.c_int32() for _ in range(4)]
        _check(lib().cgl_tuned_params(self.h, *(ctypes.byref(x) for x in v)))
        return dict(xchunk=v[0].value, xcd_order=v[1].value, tile_rows=v[2].value, tuning=bool(v[3].value))

    def scatter_form(self):
        """True when the steps sum the next phase field in scatter form (one
        read of the state per step)."""
        v = ctypes.c_int32()
        _check(lib().cgl_phase_field_form(self.h, ctypes.byref(v)))
        return bool(v.value)

    def save(self, path):
        """Checkpoint of the device state (resumes bit for bit)."""
        _check(lib().cgl_checkpoint_save(self.h, os.fsencode(path)))

    def load(self, path):
        _check(lib().cgl_checkpoint_load(self.h, os.fsencode(path)))

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        _check(lib().cgl_device_name(self.h, buf, 256))
        return buf.value.decode()


class Batch:
    """Independent periodic boxes of uneven sizes advanced together (one
    launch per kernel over all boxes, steps replayed from a HIP graph).
    ``boxes``: sequence of (nx, ny, nz, nu_r, nu_b)."""

    def __init__(self, boxes, precision="f32", sigma=0.01, rho0=1.0, beta=0.7, xchunk=0, device=0, layout="fs"):
        p = Params()
        p.precision = PREC[precision] if isinstance(precision, str) else int(precision)
        p.layout = LAYOUT[layout] if isinstance(layout, str) else int(layout)
        p.xchunk = xchunk
        p.device = device
        p.sigma, p.rho0, p.beta = sigma, rho0, beta
        arr = (Box * len(boxes))()
        for k, (nx, ny, nz, nu_r, nu_b) in enumerate(boxes):
            arr[k].nx, arr[k].ny, arr[k].nz, arr[k].nu_r, arr[k].nu_b = nx, ny, nz, nu_r, nu_b
        self.shapes = [tuple(b[:3]) for b in boxes]
        self.nodes = sum(nx * ny * nz for nx, ny, nz in self.shapes)
        self.h = ctypes.c_void_p()
        _check(lib().cgl_batch_create(ctypes.byref(self.h), ctypes.byref(p), len(boxes), arr))

    def close(self):
        if self.h:
            lib().cgl_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, b, ri, bi):
        ri = np.ascontiguousarray(ri, dtype=np.float64)
        bi = np.ascontiguousarray(bi, dtype=np.float64)
        assert ri.shape == self.shapes[b] + (Q,) and bi.shape == ri.shape
        _check(lib().cgl_box_upload_state(self.h, b, _ptr(ri), _ptr(bi)))

    def init_cylinder(self, b, xc, yc, r0, kprof, amps):
        a = np.ascontiguousarray(amps, dtype=np.float64)
        _check(lib().cgl_box_init_cylinder(self.h, b, xc, yc, r0, kprof, len(a), _ptr(a)))

    def download(self, b):
        ri = np.empty(self.shapes[b] + (Q,))
        bi = np.empty(self.shapes[b] + (Q,))
        _check(lib().cgl_box_download_state(self.h, b, _ptr(ri), _ptr(bi)))
        return ri, bi

    def calmacros(self, b):
        names = ("rho", "rho_r", "rho_b", "ux", "uy", "uz", "phi")
        out = {k: np.empty(self.shapes[b]) for k in names}
        _check(lib().cgl_box_calmacros(self.h, b, *(_ptr(out[k]) for k in names)))
        return out

    def scatter_form(self):
        v = ctypes.c_int32()
        _check(lib().cgl_phase_field_form(self.h, ctypes.byref(v)))
        return bool(v.value)

    def save(self, path):
        _check(lib().cgl_checkpoint_save(self.h, os.fsencode(path)))

    def load(self, path):
        _check(lib().cgl_checkpoint_load(self.h, os.fsencode(path)))

    def step(self, n=1):
        _check(lib().cgl_step(self.h, int(n)))

    def sync(self):
        _check(lib().cgl_sync(self.h))

    def time_steps(self, n):
        t, m, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().cgl_time_steps(self.h, int(n), ctypes.byref(t), ctypes.byref(m), ctypes.byref(c)))
        return t.value, m.value, c.value
