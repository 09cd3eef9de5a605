"""FP64 CPU oracle for the colour-gradient D3Q19 MRT lattice update.

TEST INFRASTRUCTURE ONLY.  This package is the checker: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
it.  The product (``symmetry_amd``) never imports, loads or calls anything
here.

* ``OracleState`` wraps ``liboracle.so``, the C restatement of the
  reference's stages (Source1.f90, cited as S:N in ``cgl_oracle.c``).
* ``RefHarness`` wraps ``_ref/libref.so``, the unmodified reference compiled
  by ``Makefile`` (target ``ref``) from the reference checkout, driven through
  ``ref_harness.f90``.  It only exists where the reference checkout was
  present at build time.

Parity status of each oracle function (what pins it) is listed in the
module-level ``PINS`` table and exercised by ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")
REF_O3_PATH = os.path.join(HERE, "_ref", "o3", "libref.so")  # the CPU baseline's -O3 build
REF_SOURCE = "/root/reference/Source1.f90"

BC_PERIODIC, BC_WALL, BC_MIRROR = 0, 1, 2
Q = 19

# What pins each oracle function (see tests/test_oracle_*.py).
PINS = {
    "or_tables": "closed forms + reference DATA table fixture (tests/golden/d3q19_tables.json)",
    "or_equilf/or_equilf_srt": "meq == M @ feq_srt identity; Guo moments == M @ Guo populations",
    "or_calmacros": "compiled reference at 152^3 (S:770-831); conservation",
    "or_derivatives": "compiled reference at 152^3 incl. ghost-phi and wetting (S:452-597)",
    "or_collision": "compiled reference at 152^3 (S:599-695); uniform rest state fixed point",
    "or_recolor": "compiled reference at 152^3 (S:697-732); sum r == rho_R, r+b == f",
    "or_boundary": "compiled reference at 152^3 (S:1544-1592)",
    "or_streaming": "compiled reference (S:735-768); permutation identity",
    "or_curve": "compiled reference with 12 particles (S:990-1034)",
    "or_set_walls": "compiled reference (bit-exact ints, S:326-395); brute-force sphere count",
    "or_refilling": "compiled reference (S:1351-1418)",
    "or_momentum": "compiled reference (S:1178-1261)",
    "or_lagrange_capillary": "compiled reference (S:1799-2028) at the reference's contact angle of 90 degrees; "
                             "other angles (the S:2015-2019 else-branch, which the reference never runs) by "
                             "closed forms of the force direction and size (tests/test_oracle_contact_angle.py)",
    "wetting at other angles": "closed form of the corrected normal on a linear phase field, 30-150 degrees "
                               "(tests/test_oracle_contact_angle.py)",
    "periodic faces": "translation invariance of a step (tests/test_oracle_invariance.py); the reference has none",
    "high-side wall faces": "reflection symmetry of a step between two walls (tests/test_oracle_invariance.py)",
    "or_repulsive": "compiled reference (S:1594-1671)",
    "or_movement": "compiled reference (S:1263-1349)",
    "or_step": "compiled reference, whole steps with and without particles",
    "energy.energy (output4)": "compiled reference's own output4 at 152^3 with its 12 particles: Energy.dat line, "
                               "Wcap_old and Shb (S:1719-1797)",
}


def build(force: bool = False, ref: bool | None = None) -> None:
    """Compile liboracle.so (always) and _ref/libref.so (when the reference
    checkout is present).  ``make`` skips up-to-date targets unless forced."""
    args = ["make", "-s", "-C", HERE]
    if force:
        args.append("-B")
    targets = ["all"]
    if ref is None:
        ref = os.path.exists(REF_SOURCE)
    if ref:
        targets.append("ref")
    subprocess.run(args + targets, check=True)


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def tables():
    """(wi, c[3,19], opc, sym[3,19], M[19,19], Minv[19,19]) from the oracle."""
    wi = np.zeros(Q)
    c = np.zeros((3, Q), dtype=np.int32)
    opc = np.zeros(Q, dtype=np.int32)
    sym = np.zeros((3, Q), dtype=np.int32)
    M = np.zeros((Q, Q))
    Minv = np.zeros((Q, Q))
    lib().or_tables(*(a.ctypes.data_as(ctypes.c_void_p) for a in (wi, c, opc, sym, M, Minv)))
    return wi, c, opc, sym, M, Minv


def equilf_srt(rho, ux, uy, uz):
    out = np.zeros(Q)
    f = lib().or_equilf_srt
    f.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 4
    f(out.ctypes.data, rho, ux, uy, uz)
    return out


def equilf(rho, ux, uy, uz, fx, fy, fz):
    meq = np.zeros(Q)
    fk = np.zeros(Q)
    f = lib().or_equilf
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p] + [ctypes.c_double] * 7
    f(meq.ctypes.data, fk.ctypes.data, rho, ux, uy, uz, fx, fy, fz)
    return meq, fk


_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


class _OrState(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("bc", ctypes.c_int * 6),
        ("npar", ctypes.c_int),
        ("sigma", ctypes.c_double), ("rho0", ctypes.c_double), ("nu_r", ctypes.c_double),
        ("nu_b", ctypes.c_double), ("beta", ctypes.c_double),
        ("fg", ctypes.c_double),
        ("rep_plane", ctypes.c_double * 6), ("rep_on", ctypes.c_int * 6),
        ("dom_lo", ctypes.c_double * 3), ("dom_hi", ctypes.c_double * 3),
        ("dom_lo_on", ctypes.c_int * 3), ("dom_lo_strict", ctypes.c_int * 3), ("dom_hi_on", ctypes.c_int * 3),
        ("ref_rules", ctypes.c_int),
        ("ri", _DP), ("bi", _DP), ("fi", _DP), ("ffi", _DP),
        ("rho", _DP), ("rho_r", _DP), ("rho_b", _DP), ("phi", _DP), ("ux", _DP), ("uy", _DP), ("uz", _DP),
        ("gx", _DP), ("gy", _DP), ("gz", _DP), ("nxv", _DP), ("nyv", _DP), ("nzv", _DP),
        ("flag", _IP),
        ("walls", _IP), ("walls_old", _IP), ("walls_all", _IP), ("pin", _IP), ("pout", _IP),
        ("center", _DP), ("center_old", _DP), ("center0", _DP),
        ("up", _DP), ("rp", _DP),
        ("fp", _DP), ("tp", _DP), ("frep", _DP), ("tension", _DP), ("tension_t", _DP),
        ("rad", _DP), ("mass", _DP), ("inertia", _DP), ("cita", _DP), ("shb", _DP),
    ]


POP_FIELDS = ("ri", "bi", "fi", "ffi")
SCALAR_FIELDS = ("rho", "rho_r", "rho_b", "phi", "ux", "uy", "uz", "gx", "gy", "gz", "nxv", "nyv", "nzv")
INT_FIELDS = ("flag", "walls", "walls_old", "walls_all", "pin", "pout")
PART_VEC = ("center", "center_old", "center0", "up", "rp", "fp", "tp", "frep", "tension", "tension_t")
PART_SCAL = ("rad", "mass", "inertia", "cita", "shb")


class OracleState:
    """FP64 state of the oracle.  Arrays are numpy views in the reference's
    layout with 0-based indices: scalars [x, y, z], populations [x, y, z, i]."""

    def __init__(self, nx, ny, nz, bc=(0, 0, 0, 0, 0, 0), npar=0, sigma=0.01, rho0=1.0,
                 nu_r=1.0 / 6.0, nu_b=1.0 / 6.0, beta=0.7, fg=0.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.npar = npar
        self.arrays = {}
        for k in POP_FIELDS:
            self.arrays[k] = np.zeros((nx, ny, nz, Q))
        for k in SCALAR_FIELDS:
            self.arrays[k] = np.zeros((nx, ny, nz))
        for k in INT_FIELDS:
            self.arrays[k] = np.zeros((nx, ny, nz), dtype=np.int32)
        npa = max(npar, 1)
        for k in PART_VEC:
            self.arrays[k] = np.zeros((npa, 3))
        for k in PART_SCAL:
            self.arrays[k] = np.zeros(npa)
        self.s = _OrState()
        self.s.nx, self.s.ny, self.s.nz = nx, ny, nz
        for i, v in enumerate(bc):
            self.s.bc[i] = v
        self.s.npar = npar
        self.s.sigma, self.s.rho0, self.s.nu_r, self.s.nu_b, self.s.beta = sigma, rho0, nu_r, nu_b, beta
        self.s.fg = fg
        for k, a in self.arrays.items():
            ptype = _IP if a.dtype == np.int32 else _DP
            setattr(self.s, k, a.ctypes.data_as(ptype))

    def __getattr__(self, name):
        arrays = self.__dict__.get("arrays")
        if arrays is not None and name in arrays:
            return arrays[name]
        raise AttributeError(name)

    def call(self, fn, *extra):
        getattr(lib(), "or_" + fn)(ctypes.byref(self.s), *extra)

    def streaming(self, which):
        lib().or_streaming(ctypes.byref(self.s), self.arrays[which].ctypes.data_as(_DP))

    def step(self, nsteps=1, particles=False):
        lib().or_run(ctypes.byref(self.s), int(nsteps), int(bool(particles)))


# ------------------------------------------------------------------ reference

REF_FIELDS = {
    "fi": 0, "ffi": 1, "ri": 2, "bi": 3, "rho": 4, "rho_r": 5, "rho_b": 6, "phi": 7,
    "ux": 8, "uy": 9, "uz": 10, "gx": 11, "gy": 12, "gz": 13, "nxv": 14, "nyv": 15, "nzv": 16,
    "flag": 17, "walls": 18, "walls_old": 19, "walls_all": 20, "pin": 21, "pout": 22,
}
REF_PART = {
    "center": 23, "center_old": 24, "center0": 25,
    "Fpx": 26, "Fpy": 27, "Fpz": 28, "Tpx": 29, "Tpy": 30, "Tpz": 31,
    "Frex": 32, "Frey": 33, "Frez": 34, "tension": 35, "tension_t": 36,
    "upx": 37, "upy": 38, "upz": 39, "rpx": 40, "rpy": 41, "rpz": 42, "Shb": 43,
}
REF_CALL = {
    "calmacros": 1, "derivatives": 2, "collision": 3, "recolor": 4, "boundary": 5,
    "streaming_r": 6, "streaming_b": 7, "curve": 8, "set_walls": 9, "refilling": 10,
    "momentum": 11, "lagrange": 12, "capillary": 13, "repulsive": 14, "movement": 15,
    "initialization": 16, "initia_center": 17, "tables": 18, "output4": 19, "output3": 20,
}

# The reference's geometry: x=1 and z=1 walls, mirrors at x=xl, y=1, y=yl, z=zl
# (S:806-807, S:1443-1592).
REF_BC = (BC_WALL, BC_MIRROR, BC_MIRROR, BC_MIRROR, BC_WALL, BC_MIRROR)


_SCRATCH = None


def _scratch_dir():
    global _SCRATCH
    if _SCRATCH is None:
        import tempfile
        _SCRATCH = tempfile.mkdtemp(prefix="cgl_ref_")
    return _SCRATCH


class RefHarness:
    """The compiled reference at its own compile-time box (152^3, par=12).

    All calls run in one worker thread with a large stack, because the
    reference's streaming keeps a whole population array as a stack local
    (S:740)."""

    def __init__(self, M, sigma, rho0, nu_r, nu_b, rad, rad_c, path=REF_PATH):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        os.environ.setdefault("OMP_NUM_THREADS", "1")
        self.lib = ctypes.CDLL(path)
        self.lib.ref_ptr.restype = ctypes.c_void_p
        self.lib.ref_setup.argtypes = [ctypes.c_double] * 6 + [ctypes.c_void_p]
        Mf = np.asfortranarray(np.asarray(M, dtype=np.float64))
        self._run(self.lib.ref_setup, sigma, rho0, nu_r, nu_b, rad, rad_c, Mf.ctypes.data)
        self._run(self.lib.ref_alloc)
        self.xl, self.yl, self.zl = (self.lib.ref_dims(k) for k in range(3))
        self.par = self.lib.ref_dims(3)
        self._run(self.lib.ref_call, REF_CALL["tables"])

    @staticmethod
    def _run(fn, *args):
        """Run a reference call on a large-stack thread, in a scratch working
        directory: the reference's Lagrange routine writes Laglange.dat into
        the current directory every call."""
        box = {}

        def work():
            try:
                box["r"] = fn(*args)
            except BaseException as e:  # pragma: no cover - propagated below
                box["e"] = e

        old = threading.stack_size()
        cwd = os.getcwd()
        os.chdir(_scratch_dir())
        threading.stack_size(1 << 30)
        try:
            th = threading.Thread(target=work)
            th.start()
            th.join()
        finally:
            threading.stack_size(old)
            os.chdir(cwd)
        if "e" in box:
            raise box["e"]
        return box.get("r")

    def field(self, name):
        p = self.lib.ref_ptr(REF_FIELDS[name])
        if name in ("fi", "ffi", "ri", "bi"):
            shape, ct = (self.xl, self.yl, self.zl, Q), ctypes.c_double
        elif name in INT_FIELDS:
            shape, ct = (self.xl, self.yl, self.zl), ctypes.c_int
        else:
            shape, ct = (self.xl, self.yl, self.zl), ctypes.c_double
        return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=shape)

    def part(self, name):
        p = self.lib.ref_ptr(REF_PART[name])
        if name in ("center", "center_old", "center0", "tension", "tension_t"):
            a = np.ctypeslib.as_array(ctypes.cast(p, _DP), shape=(3, self.par))
            return a.T  # (par, 3) view of Fortran (par, 3)
        return np.ctypeslib.as_array(ctypes.cast(p, _DP), shape=(self.par,))

    def scal(self):
        return np.ctypeslib.as_array(ctypes.cast(self.lib.ref_ptr(46), _DP), shape=(8,))

    def call(self, name):
        self._run(self.lib.ref_call, REF_CALL[name])

    def energy_line(self):
        """The line the last output4 call wrote to Energy.dat."""
        with open(os.path.join(_scratch_dir(), "Energy.dat")) as f:
            return f.read().splitlines()[-1]

    def hmin_lines(self):
        """The lines of hmin.dat after the last output3 call."""
        with open(os.path.join(_scratch_dir(), "hmin.dat")) as f:
            return f.read().splitlines()

    def step(self, nsteps=1, particles=False):
        self._run(self.lib.ref_step, int(nsteps), int(bool(particles)))
