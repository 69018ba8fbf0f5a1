"""Benchmark: Biot-Savart pair interactions per second per RK4 step.

Workload: a Gaussian vortex (amplitude 4 PI, width 0.05, centre 45N 0E,
minus its mean) on the level-7 cubed sphere: 98,304 active panels, 98,306
panel vertices (particles) and 32,766 divided parent panels; Omega = 2 PI,
dt = 0.005.  A step is one full RK4 step of all three target classes with
the state resident in HBM (bve_advance); pairs are counted as
4 (N_a (N_a - 1) + N_p N_a) per step -- active and vertex targets -- and the
parent-panel targets are computed but not counted.

Prints one JSON line.  The roofline block uses HIP events recorded by the
library around every pair-kernel launch (one launch per RK4 stage) at 20
flops per pair against the MI355X float64 vector peak.  The CPU baseline
times the oracle (the reference's loop restated in C, OpenMP over targets)
on a bounded sample of the same sum.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = 20.0
PEAK_FP64_TFLOPS = 78.6      # MI355X float64 vector
METRIC = "Biot-Savart pair-interactions/s per RK4 step at <=1e-10 rel err; adjoint/fwd ratio"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--level", type=int, default=7)
    p.add_argument("--dt", type=float, default=0.005)
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="target duration of the CPU-baseline sample (0 = skip)")
    p.add_argument("--plan", type=str, default="",
                   help="targets_per_lane,tile,workgroups for the pair kernel")
    return p.parse_args()


def cpu_baseline(grid, zeta, seconds):
    """Oracle BVEActiveRHS over the first M active targets (all sources)."""
    import oracle
    x, a = grid.active_x, grid.active_area
    n = grid.n_active
    threads = oracle.num_threads()
    m = max(64, threads * 8)
    t0 = time.perf_counter()
    oracle.active_rhs(x, zeta, a, 2 * np.pi, 0, m)
    dt0 = time.perf_counter() - t0
    m = int(min(n, max(m, m * seconds / max(dt0, 1e-3))))
    t0 = time.perf_counter()
    oracle.active_rhs(x, zeta, a, 2 * np.pi, 0, m)
    el = time.perf_counter() - t0
    pairs = m * (n - 1)
    return dict(value=pairs / el, unit="pair-interactions/s", cores=threads, kind="port",
                sample=f"oracle BVEActiveRHS (reference rounding order, IEEE divisions), "
                       f"{m} of {n} active targets x {n - 1} sources at level {grid.level}, "
                       f"{el:.1f} s, {threads} OpenMP threads")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU target-row sharding is a design, not built)", file=sys.stderr)
        sys.exit(2)

    from thesiscode_amd import _lib
    from thesiscode_amd import grid as gridmod
    from thesiscode_amd import initial_conditions as ic

    g = gridmod.cubed_sphere(args.level)
    za, zp, zq = ic.gauss_vortex(g, lat0=np.pi / 4, lon0=0.0, width=0.05)
    omega = ic.OMEGA_EARTH
    absv = zp + 2 * omega * g.particles_x[:, 2]
    na, npart, npp = g.n_active, g.n_particles, g.n_passive
    pairs_primary = 4.0 * (na * (na - 1.0) + npart * na)
    pairs_all = 4.0 * (na * (na - 1.0) + (npart + npp) * na)

    ctx = _lib.Context(0)
    ctx.set_omega(omega)
    if args.plan:
        ctx.set_plan(*[int(v) for v in args.plan.split(",")])
    ctx.load_panels(g.active_x, za, g.active_area, g.passive_x)
    ctx.load_particles(g.particles_x, absv)
    print(f"# level {args.level}: {na} active, {npart} particles, {npp} passive", file=sys.stderr,
          flush=True)

    ctx.advance(args.dt, args.warmup)
    ctx.synchronize()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    ctx.advance(args.dt, args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    launches, kernel_ms = ctx.profile()
    ctx.set_profiling(False)
    ms_per_step = el * 1e3 / max(args.steps, 1)
    value = pairs_primary * args.steps / el

    # accuracy of the measured path: first 64 active targets vs the oracle
    import oracle
    u_gpu = ctx.velocity(g.active_x[:64], g.active_x, za, g.active_area, self_end=64)
    u_ref, _ = oracle.active_rhs(g.active_x, za, g.active_area, omega, 0, 64)
    u_ref = u_ref[:64]
    err = float(np.max(np.linalg.norm(u_gpu - u_ref, axis=1)) / np.max(np.linalg.norm(u_ref, axis=1)))
    ctx.close()

    avg_launch_ms = kernel_ms / max(launches, 1)
    flops_per_launch = pairs_all / 4.0 * FLOPS_PER_PAIR
    achieved = flops_per_launch / (avg_launch_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"gauss_vortex_cubed_sphere_L{args.level}_rk4_step",
            "level": args.level, "n_active": na, "n_particles": npart, "n_passive": npp,
            "dt": args.dt, "precision": "fp64",
            "pairs_per_step": pairs_primary, "pairs_per_step_all_classes": pairs_all,
            "pair_rate_all_classes": pairs_all * args.steps / el,
            "rel_err_vs_fp64_oracle": err,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": None,
            "kernel": "pair_sum_f64_kernel",
            "launches": launches,
            "avg_launch_ms": avg_launch_ms,
            "flops_per_launch": flops_per_launch,
        },
    }
    if args.cpu_seconds > 0:
        out["cpu_baseline"] = cpu_baseline(g, za, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
