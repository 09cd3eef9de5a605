"""Batched boxes and graph replay on the GPU.

* A batch of uneven boxes in FP64 reproduces the oracle bit for bit, box by
  box: the batched kernels run the single-box code on offset buffers.
* FP32 factorised batches equal single-box FP32 runs bit for bit.
* Steps replayed from the captured graph equal eagerly launched steps, replay
  after replay, with host work (macroscopic downloads) in between.
"""
import numpy as np
import pytest

import symmetry_amd
from symmetry_amd import ic
from tests.helpers import oracle_from

pytestmark = pytest.mark.gpu

BOXES = [  # nx, ny, nz, r0, amp, tau
    (32, 32, 64, 7.0, 0.5, 1.0),
    (20, 24, 40, 5.0, 0.3, 0.7),
    (16, 16, 33, 4.0, 0.2, 1.1),
    (24, 20, 70, 5.5, 0.4, 0.8),
]


def _print(*a):
    print(*a, flush=True)


def _cfg(box):
    nx, ny, nz, r0, amp, tau = box
    nu = (tau - 0.5) / 3.0
    return dict(nx=nx, ny=ny, nz=nz, r0=r0, amps=(amp,), sigma=0.01, rho0=1.0, nu_r=nu, nu_b=nu, beta=0.7)


def _batch(precision, layout, xchunk=0):
    cfgs = [_cfg(b) for b in BOXES]
    bt = symmetry_amd.Batch([(c["nx"], c["ny"], c["nz"], c["nu_r"], c["nu_b"]) for c in cfgs],
                            precision=precision, layout=layout, xchunk=xchunk)
    states = []
    for k, c in enumerate(cfgs):
        ri, bi, _ = ic.workload_state(c, exact=False)
        bt.upload(k, ri, bi)
        states.append((ri, bi))
    return bt, cfgs, states


@pytest.mark.parametrize("layout", ["rb", "fs"])
def test_batch_f64_bit_exact_vs_oracle(layout):
    _print("batch f64", layout)
    bt, cfgs, states = _batch("f64", layout, xchunk=4)
    nsteps = 9
    bt.step(nsteps)
    for k, c in enumerate(cfgs):
        st = oracle_from(c, *states[k])
        st.step(nsteps)
        r2, b2 = bt.download(k)
        assert np.array_equal(r2, st.ri) and np.array_equal(b2, st.bi), k
        st.call("calmacros")
        mac = bt.calmacros(k)
        assert np.array_equal(mac["phi"], st.phi), k


@pytest.mark.parametrize("scatter", ["1", "0"])
def test_batch_f32_equals_single_boxes(monkeypatch, scatter):
    """Batched FP32 boxes equal single boxes stepped in the same form bit for
    bit: the scatter form (its sums do not depend on the tiling) and the
    gather form."""
    _print("batch f32 fs vs single, scatter", scatter)
    monkeypatch.setenv("CGL_SCATTER", scatter)
    bt, cfgs, states = _batch("f32", "fs", xchunk=8)
    assert bt.scatter_form() == (scatter == "1")
    bt.step(20)
    for k, c in enumerate(cfgs):
        lat = symmetry_amd.Lattice(c["nx"], c["ny"], c["nz"], precision="f32", sigma=c["sigma"], rho0=c["rho0"],
                                   nu_r=c["nu_r"], nu_b=c["nu_b"], beta=c["beta"], layout="fs", xchunk=2)
        lat.upload(*states[k])
        lat.step(20)
        a, b = lat.calmacros(), bt.calmacros(k)
        for name in ("rho", "ux", "phi"):
            assert np.array_equal(a[name], b[name]), (k, name)
        lat.close()


def test_batch_device_ic_matches_single_box():
    _print("batch device ic")
    cfgs = [_cfg(b) for b in BOXES]
    bt = symmetry_amd.Batch([(c["nx"], c["ny"], c["nz"], c["nu_r"], c["nu_b"]) for c in cfgs], precision="f32",
                            layout="fs")
    for k, c in enumerate(cfgs):
        bt.init_cylinder(k, c["nx"] / 2 + 0.5, c["ny"] / 2 + 0.5, c["r0"], 6.0 * 0.7 * 0.134, c["amps"])
    bt.step(5)
    c = cfgs[1]
    lat = symmetry_amd.Lattice(c["nx"], c["ny"], c["nz"], precision="f32", nu_r=c["nu_r"], nu_b=c["nu_b"],
                               layout="fs", xchunk=4)
    lat.init_cylinder(c["nx"] / 2 + 0.5, c["ny"] / 2 + 0.5, c["r0"], 6.0 * 0.7 * 0.134, c["amps"])
    lat.step(5)
    assert np.array_equal(lat.calmacros()["phi"], bt.calmacros(1)["phi"])


def test_batch_rejects_late_upload():
    bt, cfgs, states = _batch("f32", "fs")
    bt.step(1)
    with pytest.raises(symmetry_amd.CglError):
        bt.upload(0, *states[0])


@pytest.mark.parametrize("precision,layout", [("f64", "rb"), ("f32", "fs")])
def test_graph_replay_equals_eager_steps(precision, layout):
    """cgl_step replays two captured steps; cgl_time_steps launches eagerly
    (two steps per call: its total pass and its per-kernel pass)."""
    _print("graph replay", precision, layout)
    c = _cfg(BOXES[0])
    ri, bi, _ = ic.workload_state(c, exact=False)
    lats = []
    for _ in range(2):
        lat = symmetry_amd.Lattice(c["nx"], c["ny"], c["nz"], precision=precision, nu_r=c["nu_r"], nu_b=c["nu_b"],
                                   layout=layout, xchunk=4)
        lat.upload(ri, bi)
        lats.append(lat)
    graph, eager = lats
    for rep in range(5):
        graph.step(2)
        eager.time_steps(1)
        a, b = graph.calmacros(), eager.calmacros()
        for name in ("rho", "ux", "phi"):
            assert np.array_equal(a[name], b[name]), (rep, name)
    r1, b1 = graph.download()
    r2, b2 = eager.download()
    assert np.array_equal(r1, r2) and np.array_equal(b1, b2)
