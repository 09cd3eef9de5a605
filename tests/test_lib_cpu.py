"""CPU-side checks of the product library and host code (no GPU needed):
the C ABI loads and exports every symbol include/cgl.h declares, the
parameter struct layout agrees between C and the bindings, argument
validation happens before any device call, and the host initial condition
reproduces the reference's initialization."""
import ctypes
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import symmetry_amd
from symmetry_amd import ic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "cgl.h")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(symmetry_amd.LIB_PATH):
        subprocess.run(["make", "-s", "-C", os.path.dirname(symmetry_amd.LIB_PATH)], check=True)
    return symmetry_amd.lib()


def declared_functions():
    text = open(HDR).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cgl_[a-z_0-9]+)\s*\(", text)))


def test_header_symbols_exported(lib):
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(symmetry_amd.EXPORTS)


def test_params_layout_matches_c(tmp_path):
    src = tmp_path / "off.c"
    src.write_text(
        '#include <stddef.h>\n#include <stdio.h>\n#include "cgl.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(cgl_params), offsetof(cgl_params, precision),"
        " offsetof(cgl_params, device), offsetof(cgl_params, sigma), offsetof(cgl_params, beta),"
        " offsetof(cgl_params, bc));return 0;}\n")
    exe = tmp_path / "off"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    P = symmetry_amd.Params
    assert out == [ctypes.sizeof(P), P.precision.offset, P.device.offset, P.sigma.offset, P.beta.offset,
                   P.bc.offset]


def _params(**kw):
    p = symmetry_amd.Params()
    p.nx, p.ny, p.nz = 8, 8, 8
    p.precision = 0
    p.sigma, p.rho0, p.nu_r, p.nu_b, p.beta = 0.01, 1.0, 1 / 6, 1 / 6, 0.7
    for k, v in kw.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("kw,code", [({"nx": 0}, -1), ({"precision": 7}, -1), ({"npar": -1}, -1),
                                     ({"npar": 3, "layout": 2}, -1), ({"rho0": 0.0}, -1),
                                     ({"tile_rows": 3}, -1), ({"tile_rows": 32}, -1),
                                     # largest boxes: above 2^28 nodes in FP32 and 2^27 (512^3) in FP64
                                     # a group of four direction arrays exceeds the 4 GiB per-array
                                     # addressing, and so do the FP64 factorised layout's 48 B of
                                     # recolour scalars per node at 512^3; 2048 x 1024 x 1024 the node index
                                     ({"nx": 1024, "ny": 512, "nz": 513, "precision": 1}, -3),
                                     ({"nx": 512, "ny": 512, "nz": 513}, -3),
                                     ({"nx": 512, "ny": 512, "nz": 512, "layout": 1}, -3),
                                     ({"nx": 2048, "ny": 1024, "nz": 1024, "precision": 1}, -1)])
def test_create_validates_before_touching_the_device(lib, kw, code):
    h = ctypes.c_void_p()
    rc = lib.cgl_create(ctypes.byref(h), ctypes.byref(_params(**kw)))
    assert rc == code
    assert lib.cgl_last_error()


def _accepted(lib, p):
    """Validation passes: the call reaches the device (CGL_ERR_HIP where there
    is none)."""
    h = ctypes.c_void_p()
    rc = lib.cgl_create(ctypes.byref(h), ctypes.byref(p))
    if rc == 0:
        lib.cgl_destroy(h)
    return rc in (0, -2)


def test_face_validation(lib):
    h = ctypes.c_void_p()
    # particles, walls and mirrors run on both layouts
    p = _params(npar=2, layout=1)
    p.bc[4] = p.bc[5] = 2
    assert _accepted(lib, p)
    p = _params()
    p.bc[4] = symmetry_amd.BC_WALL  # a periodic face needs a periodic partner
    assert lib.cgl_create(ctypes.byref(h), ctypes.byref(p)) == -1
    p = _params(layout=1)
    p.bc[4] = p.bc[5] = symmetry_amd.BC_WALL
    assert _accepted(lib, p)
    p = _params(layout=1)
    p.bc[0] = p.bc[1] = 2
    assert _accepted(lib, p)
    p = _params(nz=3)
    p.bc[4] = p.bc[5] = 2  # a bounded axis needs ghost layers and live nodes
    assert lib.cgl_create(ctypes.byref(h), ctypes.byref(p)) == -1
    p = _params()
    p.bc[4] = 7
    assert lib.cgl_create(ctypes.byref(h), ctypes.byref(p)) == -1


@pytest.mark.skipif(not oracle.ref_available(), reason="reference build absent")
def test_host_ic_reproduces_reference_initialization():
    with open(os.path.join(ROOT, "tests", "golden", "d3q19_tables.json")) as f:
        M = np.array(json.load(f)["M"], dtype=float)
    rad_c, sigma, rho0 = 30.0, 0.05, 0.05
    nu_r = 0.1 / math.sqrt(rho0 / (rad_c * sigma))
    ref = oracle.RefHarness(M, sigma, rho0, nu_r, nu_r / 10.0, 7.5, rad_c)
    ref.scal()[3] = (ref.yl - 2) * 2.0
    ref.call("initialization")
    xl = ref.xl
    ri, bi, phi = ic.cylinder_state(xl, ref.yl, ref.zl, axis="y", center=(xl - 0.5, ref.zl - 0.5), r0=rad_c,
                                    amps=(0.1 * rad_c,), wavelength=(ref.yl - 2) * 2.0, phase0=1.5,
                                    sigma=sigma, rho0=rho0)
    assert np.array_equal(phi, ref.field("phi"))
    assert np.array_equal(ri, ref.field("ri"))
    assert np.array_equal(bi, ref.field("bi"))


def test_config3_amplitudes_fixture():
    """The 8 mode amplitudes of the 256x256x512 workload are committed data."""
    with open(os.path.join(ROOT, "tests", "golden", "rp_config3_amps.json")) as f:
        gold = json.load(f)
    a = ic.rp_config3_amps()
    assert np.array_equal(a, np.array(gold["amps"]))
    assert np.all((a >= 0) & (a <= 0.02 * 40.0))


def test_ensemble_draws_match_fixture():
    with open(os.path.join(ROOT, "tests", "golden", "ensemble_config4.json")) as f:
        fix = json.load(f)["boxes"]
    boxes = ic.ensemble_config4()
    assert len(boxes) == len(fix) == 64
    for b, g in zip(boxes, fix):
        for k in ("nx", "ny", "nz", "r0", "tau"):
            assert b[k] == g[k]
        assert list(b["amps"]) == g["amps"]
    for b in boxes:
        assert b["nx"] in (32, 48, 64, 96) and b["nz"] % 32 == 0 and 64 <= b["nz"] <= 512
        assert 0.15 * b["nx"] <= b["r0"] <= 0.3 * b["nx"] and 0.6 <= b["tau"] <= 1.2
        assert 0.01 * b["r0"] <= b["amps"][0] <= 0.1 * b["r0"]


def test_particle_draws_match_fixture():
    """The 200 spheres of the dense interface workload (seed 99) are committed
    data; the draw rules hold: no two spheres closer than 2R + 1 (minimum
    image in x and y), centres within +-24 of the interface, densities and
    contact angles in their ranges."""
    with open(os.path.join(ROOT, "tests", "golden", "particle_config5.json")) as f:
        fix = json.load(f)
    c = ic.particle_config5()
    assert np.array_equal(c["centers"], np.array(fix["centers"]))
    assert np.array_equal(c["rho_p"], np.array(fix["rho_p"])) and np.array_equal(c["cita_deg"], np.array(fix["cita_deg"]))
    n, rad, ctr = c["nx"], c["rad"], c["centers"]
    assert len(ctr) == 200 and np.all(np.abs(ctr[:, 2] - n / 2.0) <= 24.0)
    d = np.abs(ctr[:, None, :] - ctr[None, :, :])
    d[..., :2] = np.minimum(d[..., :2], n - d[..., :2])
    dist = np.sqrt((d * d).sum(-1)) + np.eye(len(ctr)) * 1e9
    assert dist.min() >= 2.0 * rad + 1.0
    assert np.all((0.8 <= c["rho_p"]) & (c["rho_p"] <= 2.0)) and np.all((30 <= c["cita_deg"]) & (c["cita_deg"] <= 150))


def test_batch_create_validates(lib):
    boxes = (symmetry_amd.Box * 2)()
    boxes[0].nx, boxes[0].ny, boxes[0].nz, boxes[0].nu_r, boxes[0].nu_b = 8, 8, 8, 0.1, 0.1
    boxes[1].nx, boxes[1].ny, boxes[1].nz, boxes[1].nu_r, boxes[1].nu_b = 0, 8, 8, 0.1, 0.1
    h = ctypes.c_void_p()
    assert lib.cgl_batch_create(ctypes.byref(h), ctypes.byref(_params()), 2, boxes) == -1
    boxes[1].nx = 8
    p = _params()
    p.bc[0] = p.bc[1] = 1
    assert lib.cgl_batch_create(ctypes.byref(h), ctypes.byref(p), 2, boxes) == -3
