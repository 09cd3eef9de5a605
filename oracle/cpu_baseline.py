"""CPU baseline timing (test infrastructure; bench.py's cpu_baseline leg).

Times the reference's own no-particle time step -- calmacros, derivatives,
collision, Recolor, boundary, streaming x2, curve (Source1.f90 S:216-251 with
the particle block switched off as its commented if(1==2) does) -- from the
unmodified reference compiled at -O3 with OpenMP into oracle/_ref/o3/libref.so
(the -O2, no-FMA checker build oracle/_ref/libref.so where that is absent),
over all requested threads.  The reference's box is its compile-time 152^3.  When
the reference build is absent, times the FP64 C restatement
(liboracle.so, one thread) on a periodic box instead.

Run as a subprocess so OMP_NUM_THREADS takes effect before libomp starts:
    python -m oracle.cpu_baseline --threads 16 --seconds 20
Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time


def cpu_model():
    """The host CPU's model name and its logical CPU count."""
    name = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return f"{name}, {os.cpu_count()} logical CPUs visible"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--seconds", type=float, default=20.0, help="approximate CPU time budget")
    ap.add_argument("--port", action="store_true", help="time the C restatement even if the reference is built")
    args = ap.parse_args()
    os.environ["OMP_NUM_THREADS"] = str(args.threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import numpy as np

    import oracle
    from symmetry_amd import ic

    ref_path = oracle.REF_O3_PATH if os.path.exists(oracle.REF_O3_PATH) else oracle.REF_PATH
    flags = "-O3" if ref_path == oracle.REF_O3_PATH else "-O2 -ffp-contract=off"
    if os.path.exists(ref_path) and not args.port:
        with open(os.path.join(oracle.HERE, "..", "tests", "golden", "d3q19_tables.json")) as f:
            M = np.array(json.load(f)["M"], dtype=float)
        rad_c, sigma, rho0 = 30.0, 0.05, 0.05
        nu_r = 0.1 / math.sqrt(rho0 / (rad_c * sigma))
        ref = oracle.RefHarness(M, sigma, rho0, nu_r, nu_r / 10.0, 7.5, rad_c, path=ref_path)
        ref.scal()[3] = (ref.yl - 2) * 2.0
        ref.part("center")[:] = -1000.0  # no particles in the box
        ref.call("set_walls")
        ref.call("initialization")
        ref.call("derivatives")
        t0 = time.perf_counter()
        ref.step(1, particles=False)  # warm-up, also sizes the sample
        t1 = time.perf_counter()
        nsteps = max(2, int(args.seconds / max(t1 - t0, 1e-3)))
        t0 = time.perf_counter()
        ref.step(nsteps, particles=False)
        dt = time.perf_counter() - t0
        nodes = ref.xl * ref.yl * ref.zl
        out = dict(value=nodes * nsteps / dt / 1e6, unit="MLUPS", cores=args.threads, kind="reference",
                   sample=f"reference Source1.f90 (flang {flags}, OpenMP) no-particle FP64 step at its compiled "
                          f"{ref.xl}x{ref.yl}x{ref.zl} box, {nsteps} steps after 1 warm-up, {dt:.1f} s, "
                          f"{args.threads} threads on {cpu_model()}")
    else:
        nx, ny, nz = 64, 64, 128
        cfg = ic.rp_config3()
        ri, bi, _ = ic.workload_state(cfg, exact=False, nx=nx, ny=ny, nz=nz)
        st = oracle.OracleState(nx, ny, nz, sigma=cfg["sigma"], rho0=cfg["rho0"], nu_r=cfg["nu_r"],
                                nu_b=cfg["nu_b"], beta=cfg["beta"])
        st.ri[...] = ri
        st.bi[...] = bi
        t0 = time.perf_counter()
        st.step(1)
        t1 = time.perf_counter()
        nsteps = max(2, int(args.seconds / max(t1 - t0, 1e-3)))
        t0 = time.perf_counter()
        st.step(nsteps)
        dt = time.perf_counter() - t0
        out = dict(value=nx * ny * nz * nsteps / dt / 1e6, unit="MLUPS", cores=1, kind="port",
                   sample=f"FP64 C restatement of the reference step, periodic {nx}x{ny}x{nz} slice of the "
                          f"256x256x512 workload, {nsteps} steps, {dt:.1f} s, one thread on {cpu_model()}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
