"""GPU: checkpoint and restart (cgl_checkpoint_save / cgl_checkpoint_load).
The reference has none; the contract is that a run stopped, saved, loaded
into a fresh handle of the same configuration and continued equals the
uninterrupted run bit for bit -- for every storage precision and layout
(including the scatter-form phase field carried between steps), for
particles with walls and in the reference's mirror box (masks, band fields,
particle state, the cold-start flag), and for batches."""
import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import small_rp
from tests.test_gpu_batch import _batch
from tests.test_gpu_ext import _lattice_particles, _particle_setup
from tests.test_gpu_mirror_particles import _lattice as _mirror_lattice
from tests.test_gpu_mirror_particles import _setup as _mirror_setup

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _fluid(cfg, precision, layout):
    lat = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                               rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], layout=layout,
                               xchunk=8)
    return lat


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("precision,layout", [("f64", "rb"), ("f32", "fs"), ("f16", "fs"), ("f32", "rb")])
def test_fluid_restart_bit_exact(tmp_path, precision, layout):
    cfg, ri, bi = small_rp(nx=32, ny=16, nz=128, nu_b=0.1)
    _print("restart", precision, layout)
    a = _fluid(cfg, precision, layout)
    a.upload(ri, bi)
    a.step(7)
    path = tmp_path / "c.bin"
    a.save(path)
    a.step(9)
    b = _fluid(cfg, precision, layout)
    b.load(path)
    b.step(9)
    assert _same(a.download(), b.download())
    ma, mb = a.calmacros(), b.calmacros()
    assert all(np.array_equal(ma[k], mb[k]) for k in ma)


@pytest.mark.parametrize("precision,layout,before", [("f64", "rb", 1), ("f64", "fs", 1), ("f32", "fs", 1),
                                                    ("f32", "fs", 4), ("f16", "fs", 4)])
def test_particles_restart_bit_exact(tmp_path, precision, layout, before):
    """Saved past the cold start; after four steps the shifted-storage states
    are in the scatter form's run (the phase-field buffer holds the last
    step's field), which the checkpoint carries."""
    cfg, parts, ri, bi = _particle_setup()
    _print("particle restart", precision, layout, before)
    a = _lattice_particles(cfg, parts, ri, bi, precision, layout=layout)
    a.step(before)
    path = tmp_path / "p.bin"
    a.save(path)
    a.step(8)
    b = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision=precision, sigma=cfg["sigma"],
                             rho0=cfg["rho0"], nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], bc=cfg["bc"],
                             npar=len(parts["rad"]), layout=layout)
    b.load(path)
    b.step(8)
    assert _same(a.download(), b.download())
    pa, pb = a.get_particles(), b.get_particles()
    assert all(np.array_equal(pa[k], pb[k]) for k in pa)
    ma, mb = a.masks(), b.masks()
    assert all(np.array_equal(ma[k], mb[k]) for k in ma)


def test_particles_restart_at_cold_start(tmp_path):
    """Saved right after set_particles: the restart keeps the start-up
    (ffi = fi, S:894) of the first step."""
    cfg, parts, ri, bi = _particle_setup()
    a = _lattice_particles(cfg, parts, ri, bi)
    path = tmp_path / "p0.bin"
    a.save(path)
    a.step(5)
    b = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision="f64", sigma=cfg["sigma"], rho0=cfg["rho0"],
                             nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], bc=cfg["bc"],
                             npar=len(parts["rad"]))
    b.load(path)
    b.step(5)
    assert _same(a.download(), b.download())
    pa, pb = a.get_particles(), b.get_particles()
    assert all(np.array_equal(pa[k], pb[k]) for k in pa)


def test_mirror_particles_restart_bit_exact(tmp_path):
    cfg, parts, ri, bi, geo = _mirror_setup()
    a = _mirror_lattice(cfg, parts, ri, bi, geo)
    a.step(3)
    path = tmp_path / "m.bin"
    a.save(path)
    a.step(6)
    b = symmetry_amd.Lattice(cfg["nx"], cfg["ny"], cfg["nz"], precision="f64", sigma=cfg["sigma"], rho0=cfg["rho0"],
                             nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], bc=cfg["bc"],
                             npar=len(parts["rad"]))
    b.load(path)  # the geometry and rules come back with the particles
    b.step(6)
    assert _same(a.download(), b.download())
    pa, pb = a.get_particles(), b.get_particles()
    assert all(np.array_equal(pa[k], pb[k]) for k in pa)


def test_batch_restart_bit_exact(tmp_path):
    a, cfgs, _ = _batch("f32", "fs", xchunk=8)
    a.step(5)
    path = tmp_path / "b.bin"
    a.save(path)
    a.step(6)
    b, _, _ = _batch("f32", "fs", xchunk=8)
    b.load(path)
    b.step(6)
    for k in range(len(cfgs)):
        assert _same(a.download(k), b.download(k)), k


def test_batch_load_rejects_other_chunk_or_boxes(tmp_path):
    """A batch's tile tables belong to its creation x chunk, and its packed
    state to its box list: a checkpoint loads only into the same batch."""
    a, cfgs, _ = _batch("f32", "fs", xchunk=8)
    a.step(3)
    path = tmp_path / "b.bin"
    a.save(path)
    other_chunk, _, _ = _batch("f32", "fs", xchunk=16)
    with pytest.raises(symmetry_amd.CglError):
        other_chunk.load(path)
    boxes = [(c["nx"], c["ny"], c["nz"], c["nu_r"], c["nu_b"]) for c in cfgs]
    permuted = symmetry_amd.Batch(boxes[1:] + boxes[:1], precision="f32", layout="fs", xchunk=8)
    with pytest.raises(symmetry_amd.CglError):
        permuted.load(path)
    same, _, _ = _batch("f32", "fs", xchunk=8)
    same.load(path)
    same.step(2)
    a.step(2)
    for k in range(len(cfgs)):
        assert _same(a.download(k), same.download(k)), k


def test_load_rejects_other_configurations(tmp_path):
    cfg, ri, bi = small_rp(nx=32, ny=16, nz=128)
    a = _fluid(cfg, "f32", "fs")
    a.upload(ri, bi)
    a.step(2)
    path = tmp_path / "c.bin"
    a.save(path)
    other = _fluid(dict(cfg, nx=24), "f32", "fs")
    with pytest.raises(symmetry_amd.CglError):
        other.load(path)
    rb = _fluid(cfg, "f32", "rb")
    with pytest.raises(symmetry_amd.CglError):
        rb.load(path)
    junk = tmp_path / "junk.bin"
    junk.write_bytes(b"not a checkpoint" * 8)
    with pytest.raises(symmetry_amd.CglError):
        a.load(junk)
    a.step(1)  # the handle is untouched by the rejected loads
