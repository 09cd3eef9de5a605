#!/usr/bin/env python3
"""Benchmark: MLUPS of the per-timestep lattice update on one MI355X.

Workload: the 256x256x512 periodic Rayleigh-Plateau box (liquid cylinder of
R=40 along z with 8 seeded cosine modes, tau_red=1.0, tau_blue=0.7,
sigma=A=0.01, beta=0.7), synthetic initial state generated on the device.
A step is one full lattice update of all 33,554,432 nodes: for the periodic
FP32/FP16 factorised box the fused step kernel (which also sums the next
phase field in scatter form) and a small phase-field pass; otherwise the
moments kernel and the fused collide/stream kernel.  The headline precision is FP32 shifted
storage; --precision selects f64 / f16.

Prints ONE JSON line (rank 0).  Fields beyond the driver contract:
  roofline      the dominant kernel (fused collide) against the 8 TB/s HBM
                peak, with its algorithmic bytes per node (every stored word
                the kernel needs read once and written once)
  cpu_baseline  the reference's own FP64 step on the host cores
  step_roofline MLUPS x counted bytes per node (312 B FP32 / 624 B FP64 /
                160 B FP16) / 8 TB/s
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X
# Counted bytes per node per step (every red/blue population read and written
# once, phi written and read once).
B_ALG_STEP = {"f64": 624, "f32": 312, "f16": 160}
# Algorithmic bytes per node of the fused collide kernel alone.
#  rb: 38 populations read + 38 written + phi read once.
#  fs: 19 total populations and the per-node recolour scalars (4 x FP32, or
#      6 x FP64) read and written, phi read once.
B_ALG_COLLIDE = {
    ("f64", "rb"): 2 * 19 * 8 * 2 + 8, ("f32", "rb"): 2 * 19 * 4 * 2 + 4, ("f16", "rb"): 2 * 19 * 2 * 2 + 4,
    ("f64", "fs"): 2 * (19 * 8 + 48) + 8, ("f32", "fs"): 2 * (19 * 4 + 16) + 4, ("f16", "fs"): 2 * (19 * 2 + 16) + 4,
}
# Scatter form (periodic FP32/FP16 factorised boxes): 19 total populations read
# and written, phi read once, the node's own fixed-point sums (2 x int32)
# written once; the recolour scalars are written on a run's last step only.
B_ALG_COLLIDE_SC = {"f32": 2 * 19 * 4 + 4 + 8, "f16": 2 * 19 * 2 + 4 + 8}
# With particles the scatter step's tile kernel also reads each node's type
# (4 B); its finishing pass reads the node's own sums and writes phi.
B_ALG_COLLIDE_SC_EXT = {k: v + 4 for k, v in B_ALG_COLLIDE_SC.items()}
B_ALG_PHI_SC = 8 + 4
# Moments kernel of the red/blue layout: 38 populations read and phi written;
# shifted storage also reads the previous phi (the shift reference).
B_ALG_MOMENTS_RB = {"f64": 2 * 19 * 8 + 8, "f32": 2 * 19 * 4 + 2 * 4, "f16": 2 * 19 * 2 + 2 * 4}
# Moments kernel of the factorised layout: 19 totals and the recolour scalars
# read, phi written.
B_ALG_MOMENTS_FS = {"f64": 19 * 8 + 48 + 8, "f32": 19 * 4 + 16 + 4, "f16": 19 * 2 + 16 + 4}
DTYPE = {"f64": "f64", "f32": "f32", "f16": "f16-storage/f32-compute"}


def finish_tuning(lat, max_calls=64):
    """Run the launch-shape autotuner (x chunk, XCD tile order, tile height)
    to its end outside any timed region (it times each candidate on real
    steps), then four steady steps so that the step graph is captured before
    the clock starts."""
    for _ in range(max_calls):
        if not lat.tuned_params()["tuning"]:
            break
        lat.step(2)
    if lat.tuned_params()["tuning"]:
        raise RuntimeError("launch-shape autotuner still running after the warm-up")
    lat.step(4)
    lat.sync()


def cpu_baseline(seconds):
    threads = int(os.environ.get("CGL_CPU_THREADS", "0")) or min(os.cpu_count() or 1, 16)
    try:
        r = subprocess.run([sys.executable, "-m", "oracle.cpu_baseline", "--threads", str(threads), "--seconds",
                            str(seconds)], cwd=ROOT, capture_output=True, text=True, timeout=600)
        line = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode != 0 or not line:
            return {"value": None, "error": (r.stderr or r.stdout)[-300:]}
        return json.loads(line[-1])
    except Exception as e:  # the baseline is a report, never the measurement
        return {"value": None, "error": str(e)}


def traffic_from_profile(precision, layout, kernel="k_collide"):
    """HBM bytes per launch of the fused kernel from committed rocprofv3 PMC
    summaries (profiles/), scaled by the copy-kernel calibration; None if
    absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{precision}_{layout}_256x256x512_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("bytes_per_launch")


def _lattice(nx, ny, nz, cfg, prec, layout, scatter=None):
    saved = os.environ.get("CGL_SCATTER")
    if scatter is not None:
        os.environ["CGL_SCATTER"] = "1" if scatter else "0"
    try:
        import symmetry_amd
        lat = symmetry_amd.Lattice(nx, ny, nz, precision=prec, sigma=cfg["sigma"], rho0=cfg["rho0"],
                                   nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], layout=layout)
    finally:
        if saved is None:
            os.environ.pop("CGL_SCATTER", None)
        else:
            os.environ["CGL_SCATTER"] = saved
    lat.init_cylinder(nx / 2 + 0.5, ny / 2 + 0.5, cfg["r0"], 6.0 * cfg["beta"] * 0.134, cfg["amps"])
    return lat


def collide_kernel(lat, prec, layout):
    """(name, algorithmic bytes per node) of the step's dominant kernel."""
    if lat.scatter_form():
        return "k_collide_sc", B_ALG_COLLIDE_SC[prec]
    return "k_collide", B_ALG_COLLIDE[(prec, layout)]


def time_variant(nx, ny, nz, cfg, prec, layout, scatter, steps):
    """MLUPS of one precision/layout on the bench box (host clock around
    `steps` graph-replayed steps), and its dominant kernel's average launch
    time (HIP events on the library's stream) against the HBM roofline."""
    v = _lattice(nx, ny, nz, cfg, prec, layout, scatter)
    v.step(4)
    finish_tuning(v)
    t1 = time.perf_counter()
    v.step(steps)
    v.sync()
    nodes = nx * ny * nz
    mlups = nodes * steps / (time.perf_counter() - t1) / 1e6
    k = max(10, min(steps, 50))
    _, _, ms_col = v.time_steps(k)
    name, b = collide_kernel(v, prec, layout)
    v.close()
    avg = ms_col / k * 1e-3
    return round(mlups, 1), {"kernel": name, "alg_bytes_per_node": b, "avg_launch_ms": round(avg * 1e3, 4),
                             "frac": round(b * nodes / avg / HBM_PEAK, 4)}


def run_accuracy(nx, ny, nz, cfg, steps):
    """The bench box advanced `steps` steps from the same device initial state
    in plain FP64 (red/blue), FP32 and FP16 shifted storage (factorised,
    scatter-form phase field): L2- and Linf-relative errors of rho - rho0,
    phi and u against FP64."""
    import numpy as np

    def fields(prec, layout):
        lat = _lattice(nx, ny, nz, cfg, prec, layout)
        lat.step(steps)
        m = lat.calmacros()
        lat.close()
        return {"rho": m["rho"] - cfg["rho0"], "phi": m["phi"], "u": np.stack([m["ux"], m["uy"], m["uz"]])}

    def rel(a, b):
        d = a - b
        return [float(np.linalg.norm(d) / np.linalg.norm(b)), float(np.max(np.abs(d)) / np.max(np.abs(b)))]

    ref = fields("f64", "rb")
    out = {"steps": steps, "reference": "f64-rb", "norms": "[L2-relative, Linf-relative]"}
    for prec in ("f32", "f16"):
        m = fields(prec, "fs")
        out[f"{prec}-fs"] = {k: rel(m[k], ref[k]) for k in ("rho", "phi", "u")}
        del m
    return out


def run_ensemble(steps, seq_steps):
    """The batched ensemble: 64 uneven periodic Rayleigh-Plateau boxes (seed 7)
    in FP32 factorised storage, one launch per kernel over all boxes, steps
    replayed from a HIP graph; against the same boxes run one after another as
    plain FP64 launches (one handle per box, every step's kernels launched
    directly).  Aggregate MLUPS = total nodes x steps / wall seconds."""
    import symmetry_amd
    from symmetry_amd import ic

    boxes = ic.ensemble_config4()
    kprof = 6.0 * 0.7 * 0.134
    bt = symmetry_amd.Batch([(b["nx"], b["ny"], b["nz"], b["nu_r"], b["nu_b"]) for b in boxes], precision="f32",
                            layout="fs", sigma=0.01, rho0=1.0, beta=0.7, xchunk=8)
    for k, b in enumerate(boxes):
        bt.init_cylinder(k, b["nx"] / 2 + 0.5, b["ny"] / 2 + 0.5, b["r0"], kprof, b["amps"])
    bt.step(10)
    bt.sync()
    t0 = time.perf_counter()
    bt.step(steps)
    bt.sync()
    t_batch = time.perf_counter() - t0
    nodes = bt.nodes
    bt.close()
    # sequential plain FP64 baseline
    os.environ["CGL_GRAPHS"] = "0"
    t_seq = 0.0
    for b in boxes:
        lat = symmetry_amd.Lattice(b["nx"], b["ny"], b["nz"], precision="f64", sigma=0.01, rho0=1.0, nu_r=b["nu_r"],
                                   nu_b=b["nu_b"], beta=0.7, layout="rb")
        lat.init_cylinder(b["nx"] / 2 + 0.5, b["ny"] / 2 + 0.5, b["r0"], kprof, b["amps"])
        lat.step(14)  # x-chunk autotuning + warm-up
        lat.sync()
        t1 = time.perf_counter()
        lat.step(seq_steps)
        lat.sync()
        t_seq += time.perf_counter() - t1
        lat.close()
    del os.environ["CGL_GRAPHS"]
    batch_mlups = nodes * steps / t_batch / 1e6
    seq_mlups = nodes * seq_steps / t_seq / 1e6
    return {"workload": "64 periodic Rayleigh-Plateau boxes, nx=ny in {32,48,64,96}, nz in 64..512, seed 7",
            "boxes": len(boxes), "nodes": nodes, "batched": "f32 factorised, scatter-form phase field, one launch per kernel, graph replay",
            "steps": steps, "mlups": round(batch_mlups, 1), "sequential_f64_steps": seq_steps,
            "sequential_f64_mlups": round(seq_mlups, 1), "ratio": round(batch_mlups / seq_mlups, 2)}


def run_particles(steps, f64_dense_steps, which=("sphere_96", "dense_512"), layouts=("fs", "rb")):
    """The particle workloads: one sphere (R=10, density 1.5, 60 degrees) at a
    flat interface in a 96^3 box with z walls, FP64 (red/blue layout) and FP32
    (factorised and red/blue layouts); and 200 spheres (R=8, seed 99,
    densities in [0.8, 2.0], contact angles in [30, 150] degrees) within +-24
    of the interface in a 512^3 box with z walls, FP32 fields with FP64
    particle dynamics (factorised and red/blue), FP16 storage (factorised)
    and FP64 (red/blue).  MLUPS
    over all nodes, particle kernels included; per workload the step against
    its counted bytes (312 B/node FP32, 624 FP64), and the step's collide
    kernels and moments kernels (HIP events on the library's stream: in the
    factorised layout the scatter step's tile kernel plus the per-node kernel
    of the particles' band, and its finishing pass, which forms the next phase
    field) against the HBM roofline at the tile kernels' algorithmic bytes;
    `other_kernels_ms` is the rest of the step (particle, band-field and wall
    kernels)."""
    import symmetry_amd
    from symmetry_amd import ic

    out = {}
    c2 = ic.particle_config2()
    p2 = ic.sphere_particles([c2["center"]], c2["rad"], c2["rho_p"], c2["cita_deg"], c2["sigma"])
    c5 = ic.particle_config5()
    p5 = ic.sphere_particles(c5["centers"], c5["rad"], c5["rho_p"], c5["cita_deg"], c5["sigma"])
    runs = [("sphere_96", c2, p2, "f64", "rb", steps), ("sphere_96", c2, p2, "f32", "fs", steps),
            ("sphere_96", c2, p2, "f32", "rb", steps), ("sphere_96", c2, p2, "f16", "fs", steps),
            ("dense_512", c5, p5, "f32", "fs", steps), ("dense_512", c5, p5, "f32", "rb", steps),
            ("dense_512", c5, p5, "f16", "fs", steps)]
    if f64_dense_steps > 0:
        runs.append(("dense_512", c5, p5, "f64", "rb", f64_dense_steps))
    runs = [r for r in runs if r[0] in which and r[4] in layouts]
    roof = {}
    for name, cfg, parts, prec, layout, nsteps in runs:
        n = cfg["nx"]
        lat = symmetry_amd.Lattice(n, n, n, precision=prec, sigma=cfg["sigma"], rho0=cfg["rho0"],
                                   nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], bc=cfg["bc"],
                                   npar=len(parts["rad"]), layout=layout)
        lat.init_interface(cfg["z0"], 6.0 * cfg["beta"] * 0.134)
        lat.set_particles(parts["center"], parts["rad"], parts["mass"], parts["inertia"], parts["cita"],
                          parts["fg"])
        lat.step(20)
        finish_tuning(lat)
        t0 = time.perf_counter()
        lat.step(nsteps)
        lat.sync()
        dt = time.perf_counter() - t0
        nodes = n ** 3
        mlups = nodes * nsteps / dt / 1e6
        out[f"{name}_{prec}_{layout}_mlups"] = round(mlups, 1)
        k = 10
        _, ms_mom, ms_col = lat.time_steps(k)
        sc = lat.scatter_form()
        tuned = {key: v for key, v in lat.tuned_params().items() if key != "tuning"}
        lat.close()
        if sc:
            b_col, b_mom = B_ALG_COLLIDE_SC_EXT[prec], B_ALG_PHI_SC
            k_col = "k_collide_sc_ext (tile kernel) + k_collide_sc_band (band)"
            k_mom = "k_phi_sc4 (finishing pass of the scatter form)"
        elif layout == "rb":
            b_col, b_mom = B_ALG_COLLIDE[(prec, layout)], B_ALG_MOMENTS_RB[prec]
            k_col, k_mom = "k_collide (walls/particles)", "k_moments (walls/particles)"
        else:
            b_col, b_mom = B_ALG_COLLIDE[(prec, layout)], B_ALG_MOMENTS_FS[prec]
            k_col = "k_collide (tile kernel) + k_collide_special (band)"
            k_mom = "k_moments_fsx_ext (tile kernel) + k_moments_special (band)"
        col, mom = ms_col / k * 1e-3, ms_mom / k * 1e-3
        roof[f"{name}_{prec}_{layout}"] = {
            "steps": nsteps, "step_ms": round(dt / nsteps * 1e3, 4),
            "step_frac": round(mlups * 1e6 * B_ALG_STEP[prec] / HBM_PEAK, 4),
            "phase_field": "scatter form" if sc else "moments pass", "launch": tuned,
            "collide": {"kernel": k_col,
                        "alg_bytes_per_node": b_col,
                        "avg_launch_ms": round(col * 1e3, 4), "frac": round(b_col * nodes / col / HBM_PEAK, 4)},
            "moments": {"kernel": k_mom,
                        "alg_bytes_per_node": b_mom,
                        "avg_launch_ms": round(mom * 1e3, 4), "frac": round(b_mom * nodes / mom / HBM_PEAK, 4)},
            "other_kernels_ms": round(dt / nsteps * 1e3 - (col + mom) * 1e3, 4)}
    out["roofline"] = roof
    out["workloads"] = {"sphere_96": "96^3, z walls, one sphere R=10 at the interface",
                        "dense_512": "512^3, z walls, 200 spheres R=8 within +-24 of the interface, seed 99"}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", default="f32", choices=["f64", "f32", "f16"])
    ap.add_argument("--variant-steps", type=int, default=50,
                    help="timed steps for each other precision/layout variant (0 = skip)")
    ap.add_argument("--xchunk", type=int, default=0)
    ap.add_argument("--layout", default="fs", choices=["rb", "fs"])
    ap.add_argument("--size", default="256x256x512")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--ensemble-steps", type=int, default=500,
                    help="timed steps of the 64-box batched ensemble (0 = skip)")
    ap.add_argument("--ensemble-seq-steps", type=int, default=100,
                    help="timed steps per box of the sequential FP64 baseline")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="steps of the FP64 / FP32 / FP16 accuracy comparison on the bench box (0 = skip)")
    ap.add_argument("--particle-steps", type=int, default=200,
                    help="timed steps of the particle workloads (0 = skip)")
    ap.add_argument("--particle-f64-dense-steps", type=int, default=40,
                    help="timed steps of the 512^3 200-sphere workload in FP64 (0 = skip)")
    ap.add_argument("--particle-workloads", default="sphere_96,dense_512",
                    help="comma-separated particle workloads to run (sphere_96, dense_512)")
    ap.add_argument("--particle-layouts", default="fs,rb",
                    help="comma-separated layouts of the particle workloads (fs, rb)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this implementation runs on one GPU only; sharding over GPUs is designed, "
              "not built. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import symmetry_amd
    from symmetry_amd import ic

    nx, ny, nz = (int(v) for v in args.size.split("x"))
    cfg = ic.rp_config3()
    lat = symmetry_amd.Lattice(nx, ny, nz, precision=args.precision, sigma=cfg["sigma"], rho0=cfg["rho0"],
                               nu_r=cfg["nu_r"], nu_b=cfg["nu_b"], beta=cfg["beta"], xchunk=args.xchunk,
                               layout=args.layout)
    print(f"# device {lat.device_name()} box {nx}x{ny}x{nz} precision {args.precision} layout {args.layout}",
          file=sys.stderr, flush=True)
    copy_gbps = symmetry_amd.copy_bandwidth(2 << 30, 10)
    lat.init_cylinder(nx / 2 + 0.5, ny / 2 + 0.5, cfg["r0"], 6.0 * cfg["beta"] * 0.134, cfg["amps"])
    lat.step(args.warmup)  # with --xchunk 0 the first warm-up steps also autotune the launch shape
    finish_tuning(lat)
    tuned = lat.tuned_params()
    nodes = nx * ny * nz
    # timed region: exactly `steps` steps, synchronised on both sides
    t0 = time.perf_counter()
    lat.step(args.steps)
    lat.sync()
    wall = time.perf_counter() - t0
    mlups = nodes * args.steps / wall / 1e6
    # per-kernel durations with HIP events on the library's stream (separate pass)
    ksteps = max(10, min(args.steps, 50))
    ev_total, ms_mom, ms_col = lat.time_steps(ksteps)
    avg_col = ms_col / ksteps * 1e-3
    avg_mom = ms_mom / ksteps * 1e-3
    scatter = lat.scatter_form()
    if scatter:
        b_col = B_ALG_COLLIDE_SC[args.precision]
        kname = "k_collide_sc"
        kdesc = "k_collide_sc (fused derivatives+collision+recolour+streaming, next phase field summed in scatter form)"
    else:
        b_col = B_ALG_COLLIDE[(args.precision, args.layout)]
        kname = "k_collide"
        kdesc = "k_collide (fused derivatives+collision+recolour+streaming)"
    achieved = b_col * nodes / avg_col
    traffic = traffic_from_profile(args.precision, args.layout, kname)
    lat.close()

    # the same workload in every precision and layout: MLUPS, and the
    # dominant kernel against the HBM roofline at its algorithmic bytes
    variants, variants_roofline = {}, {}
    if args.variant_steps > 0:
        keys = [(p, l, None) for p in ("f64", "f32", "f16") for l in ("rb", "fs")]
        if scatter:
            keys.append((args.precision, args.layout, False))  # a moments pass per step (gather form)
        for prec, lay, sc in keys:
            key = f"{prec}-{lay}" + ("-gather" if sc is False else "")
            if prec == args.precision and lay == args.layout and sc is None:
                variants[key] = round(mlups, 1)
                variants_roofline[key] = {"kernel": kname, "alg_bytes_per_node": b_col,
                                          "avg_launch_ms": round(avg_col * 1e3, 4),
                                          "frac": round(achieved / HBM_PEAK, 4)}
                continue
            variants[key], variants_roofline[key] = time_variant(nx, ny, nz, cfg, prec, lay, sc, args.variant_steps)

    accuracy = run_accuracy(nx, ny, nz, cfg, args.accuracy_steps) if args.accuracy_steps > 0 else None
    ensemble = run_ensemble(args.ensemble_steps, args.ensemble_seq_steps) if args.ensemble_steps > 0 else None
    particles = (run_particles(args.particle_steps, args.particle_f64_dense_steps, args.particle_workloads.split(","),
                               args.particle_layouts.split(","))
                 if args.particle_steps > 0 else None)

    out = {
        "metric": "MLUPS at 256x256x512 D3Q19 CG-MRT, FP32 vs FP64, % of HBM roofline",
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (seeded liquid-cylinder initial state generated on the device)",
        "config": {"workload": f"Rayleigh-Plateau {nx}x{ny}x{nz} periodic D3Q19 colour-gradient MRT, R=40, "
                               "8 modes seed 1234, tau_r=1.0 tau_b=0.7, sigma=0.01, beta=0.7",
                   "precision": args.precision, "layout": args.layout, "nodes": nodes,
                   "xchunk": tuned["xchunk"], "xcd_order": tuned["xcd_order"], "tile_rows": tuned["tile_rows"],
                   "launch_autotuned": args.xchunk == 0 and not tuned["tuning"],
                   "phase_field": "scatter" if scatter else "gather"},
        "roofline": {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "kernel": kdesc,
                     "alg_bytes_per_node": b_col,
                     "avg_launch_ms": round(avg_col * 1e3, 4), "moments_avg_launch_ms": round(avg_mom * 1e3, 4)},
        "step_roofline": {"bytes_per_node": B_ALG_STEP[args.precision],
                          "frac": round(mlups * 1e6 * B_ALG_STEP[args.precision] / HBM_PEAK, 4),
                          "event_mlups": round(nodes * ksteps / (ev_total * 1e-3) / 1e6, 1)},
        "hbm_copy_gbps": round(copy_gbps, 1),
        "variants_mlups": variants,
        "variants_roofline": variants_roofline,
    }
    if accuracy:
        out["accuracy"] = accuracy
    if ensemble:
        out["ensemble"] = ensemble
    if particles:
        out["particles"] = particles
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
