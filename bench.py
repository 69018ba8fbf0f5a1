"""Benchmark: Biot-Savart pair interactions per second per RK4 step.

Workload: a Gaussian vortex (amplitude 4 PI, width 0.05, centre 45N 0E,
minus its mean) on the level-7 cubed sphere: 98,304 active panels, 98,306
panel vertices (particles) and 32,766 divided parent panels; Omega = 2 PI,
dt = 0.005.  A step is one full RK4 step of all three target classes with
the state resident in HBM (bve_advance); pairs are counted as
4 (N_a (N_a - 1) + N_p N_a) per step -- active and vertex targets -- and the
parent-panel targets are computed but not counted.

Each arithmetic mode is timed separately from the initial state; its
normwise velocity error is measured against the oracle (float64 in the
reference's rounding order) on ~190 targets of each class (active panels,
parent panels, vertices: the first 64, the 64 nearest the vortex centre, 64
random), with the resident state's velocities evaluated by the context's
own plan -- the kernel instantiation the timed steps run.  ``value`` is the
pair rate of the fastest mode whose error is <= 1e-10.

Also reported: the adjoint gradient on level 6 (cost against the forward
solve, finite-difference check), a 64-member ensemble in one HIP graph and
a level-8 two-vortex single step in every arithmetic mode.

Prints one JSON line.  The roofline block uses HIP events recorded by the
library around every pair-kernel launch (one launch per RK4 stage) at 20
flops per pair against the MI355X vector peak of the pipe the mode's pair
loop runs on (float64 78.6, float32 157.3 TFLOP/s).  The CPU baseline times
the reference's own BVERK4 (compiled unmodified into oracle/_ref/ref_bench,
MPI over 16 host cores) on one level-7 step, with three level-6 steps
beside it; the oracle (the reference's loop restated in C, OpenMP over
targets) is timed too (config.cpu_port) and stands in when the reference
build is absent.
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
PEAK_FP32_TFLOPS = 157.3     # MI355X float32 vector
ERR_BAR = 1e-10
METRIC = "Biot-Savart pair-interactions/s per RK4 step at <=1e-10 rel err; adjoint/fwd ratio"
MODES = {"fp64": 0, "fp64_fast": 2, "mixed": 1, "mixed_kahan": 3, "reference": 4,
         "mixed_bounded": 5}
# HBM bytes per launch measured with rocprofv3 FETCH_SIZE (x 2.00) + WRITE_SIZE
# (x 1.00), the factors a plain copy of known size gives (thesiscode_amd/
# copy_probe), for the default plans on the level-7 Gaussian vortex
# (profiles/pair_sum_bounded_L7_counters.md, pair_sum_f64_fast_L7_traffic_pmc.md)
MEASURED_TRAFFIC = {
    ("pair_sum_bounded_seg_kernel<2,true>", 7): 0.688e9,
    ("pair_sum_f64_seg_kernel<4,512,4>", 7): 1.99e8,
}


# Sources up to which the bounded pass reads its float32 records through
# scalar loads (kBoundedScalarMaxSources), and the source padding.
BOUNDED_SCALAR_MAX_SOURCES = 131072
SOURCE_ALIGN = 512


def kernel_name(mode, T, tile, recip, segments, n_sources=0):
    """Name of the pair kernel instantiation a mode's launch runs."""
    if mode == "mixed_bounded":
        if segments <= 0:
            return f"pair_sum_bounded_kernel<{T}>"
        n_pad = -(-n_sources // SOURCE_ALIGN) * SOURCE_ALIGN
        return f"pair_sum_bounded_seg_kernel<{T},{str(n_pad <= BOUNDED_SCALAR_MAX_SOURCES).lower()}>"
    if mode.startswith("mixed"):
        kind = "pair_sum_mixed_seg_kernel" if segments > 0 else "pair_sum_mixed_kernel"
        return f"{kind}<{T},{str(mode == 'mixed_kahan').lower()}>"
    if mode == "reference":
        return "pair_sum_exact_kernel"
    kind = "pair_sum_f64_seg_kernel" if segments > 0 else "pair_sum_f64_kernel"
    return f"{kind}<{T},{tile},{recip}>"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--level", type=int, default=7)
    p.add_argument("--dt", type=float, default=0.005)
    p.add_argument("--modes", type=str, default="fp64,fp64_fast,mixed,mixed_bounded,reference")
    p.add_argument("--cpu-seconds", type=float, default=10.0,
                   help="target duration of the CPU-baseline sample (0 = skip)")
    p.add_argument("--plan", type=str, default="",
                   help="targets_per_lane,tile,workgroups,recip for the pair kernel")
    p.add_argument("--adjoint-level", type=int, default=6,
                   help="cubed-sphere level of the adjoint workload (0 = skip)")
    p.add_argument("--adjoint-fd-panels", type=int, default=32,
                   help="panels of the adjoint's central-difference check (0 = skip it)")
    p.add_argument("--ensemble-members", type=int, default=64,
                   help="members of the ensemble workload (0 = skip)")
    p.add_argument("--large-level", type=int, default=8,
                   help="cubed-sphere level of the two-vortex large solve (0 = skip)")
    p.add_argument("--large-modes", type=str, default="fp64,fp64_fast,mixed,mixed_kahan,mixed_bounded",
                   help="arithmetic modes of the large solve")
    p.add_argument("--ensemble-modes", type=str, default="mixed,fp64_fast",
                   help="arithmetic modes of the ensemble workload")
    return p.parse_args()


def cpu_baseline(grid, zeta, seconds):
    """Oracle pair sums of whole RK4 stages (all three target classes) until
    about ``seconds`` have passed; at least one stage."""
    import oracle
    x, a = grid.active_x, grid.active_area
    n = grid.n_active
    threads = oracle.num_threads()
    pairs_stage = n * (n - 1.0) + (grid.n_particles + grid.n_passive) * n
    stages, el = 0, 0.0
    t0 = time.perf_counter()
    while stages == 0 or el + el / stages <= 1.5 * seconds:
        oracle.active_rhs(x, zeta, a, 2 * np.pi)
        oracle.passive_rhs(grid.particles_x, x, zeta, a)
        oracle.smooth_velocity(grid.passive_x, x, zeta, a, 0.05)
        stages += 1
        el = time.perf_counter() - t0
    model, _ = cpu_model()
    return dict(value=stages * pairs_stage / el, unit="pair-interactions/s", cores=threads,
                kind="port", cpu=model,
                sample=f"{stages} RK4 stage(s) of the oracle's BVEActiveRHS + BVEPassiveRHS + "
                       f"BVESmoothVelocity (reference rounding order, IEEE divisions) at level "
                       f"{grid.level}: {stages * pairs_stage:.4g} pairs in {el:.1f} s, "
                       f"{threads} OpenMP threads on {model}")


def cpu_model():
    """The host CPU's model name (first /proc/cpuinfo entry) and its logical CPUs."""
    name = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count() or 1


def cpu_baseline_reference(level=7, steps=1, ranks=16, timeout=300):
    """The reference's own BVERK4 (PanelVelocityParallel.f90, compiled
    unmodified with flang -O2 into oracle/_ref/ref_bench) timed with
    MPI_Wtime on `ranks` MPI ranks of the host (its native parallelisation:
    contiguous target ranges per rank, MPI_BCAST of every stage).  Returns
    None when the binary or mpiexec is missing."""
    import shutil
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_bench")
    mpiexec = shutil.which("mpiexec") or "/opt/conda/bin/mpiexec"
    if not (os.path.exists(exe) and os.path.exists(mpiexec)):
        return None
    ranks = max(1, min(ranks, os.cpu_count() or 1))
    try:
        r = subprocess.run([mpiexec, "-n", str(ranks), exe, str(level), str(steps)],
                           capture_output=True, text=True, timeout=timeout)
    except subprocess.TimeoutExpired:
        return None
    line = [ln for ln in r.stdout.splitlines() if ln.startswith("ref_bench")]
    if r.returncode != 0 or not line:
        return None
    f = line[-1].split()
    n_a, n_p, n_pp, nprocs, nsteps = (int(v) for v in f[2:7])
    seconds, pairs = float(f[7]), float(f[8])
    # the metric's pairs (active and vertex targets) per second of the steps
    primary = 4.0 * nsteps * (n_a * (n_a - 1.0) + n_p * n_a)
    model, logical = cpu_model()
    return dict(value=primary / seconds, unit="pair-interactions/s", cores=nprocs,
                kind="reference", cpu=model, level=level,
                sample=f"{nsteps} BVERK4 step(s) of the reference's PanelVelocityParallelModule "
                       f"(flang -O2, MPICH, {nprocs} MPI ranks, one per core, on {model}, "
                       f"{logical} logical CPUs visible) on the level-{level} cubed sphere "
                       f"({n_a} active panels, {n_p} vertices, {n_pp} parents): {pairs:.4g} "
                       f"pairs of all classes in {seconds:.1f} s; all-class rate "
                       f"{pairs / seconds:.4g}/s")


def adjoint_measurement(level, steps=20, every=5, dt=0.01, n_fd=32, mode="fp64"):
    """Gradient workload: RH4 on the level-6 cubed sphere, 20 RK4 steps at
    dt = 0.01, checkpoints every 5, J = sum_i A_i zeta_i(T) g(x_i(T)) with a
    width-0.2 Gaussian weight at 30N 90E.  Times the gradient (forward sweep
    with checkpoints + reverse sweep) against the forward solve of the same
    steps on the active panels alone (the map J depends on), with vertices,
    and with all three target classes; times it again with every stage
    input kept (no recomputation), and checks dJ/dzeta_j(0) against central
    differences of the GPU forward map on n_fd seeded panels."""
    import math
    from thesiscode_amd import _lib
    from thesiscode_amd import grid as gridmod
    from thesiscode_amd import initial_conditions as ic
    from thesiscode_amd.adjoint import GaussianWeightFunctional, central_difference, gradient

    g = gridmod.cubed_sphere(level)
    z = ic.rh4_wave(g.active_x)
    omega = ic.OMEGA_EARTH
    absv = ic.rh4_wave(g.particles_x) + 2 * omega * g.particles_x[:, 2]
    F = GaussianWeightFunctional()
    ctx = _lib.Context(0)
    ctx.set_omega(omega)
    ctx.set_precision(MODES[mode])

    def load(zz, particles=True):
        ctx.reset()
        ctx.load_panels(g.active_x, zz, g.active_area, g.passive_x)
        if particles:
            ctx.load_particles(g.particles_x, absv)

    def forward_seconds(passive, particles):
        """Wall time of `steps` resident RK4 steps (after one warm-up step)
        with the given target classes beside the active panels."""
        ctx.reset()
        ctx.load_panels(g.active_x, z, g.active_area, g.passive_x if passive else np.zeros((0, 3)))
        if particles:
            ctx.load_particles(g.particles_x, absv)
        ctx.advance(dt, 1)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.advance(dt, steps)
        ctx.synchronize()
        return time.perf_counter() - t0

    t_fwd = forward_seconds(True, True)             # all three target classes
    t_fwd_vertices = forward_seconds(False, True)   # active panels + vertices
    t_fwd_active = forward_seconds(False, False)    # the map J depends on
    load(z, particles=False)
    gradient(ctx, F, g.active_area, dt, 2, 1)
    t0 = time.perf_counter()
    J, gz, gx = gradient(ctx, F, g.active_area, dt, steps, every)
    t_adj = time.perf_counter() - t0
    # every stage input kept from the forward sweep (one checkpoint): no
    # recomputation, 56 bytes per panel and stage resident in HBM
    gradient(ctx, F, g.active_area, dt, steps, steps)
    t0 = time.perf_counter()
    J1, gz1, _ = gradient(ctx, F, g.active_area, dt, steps, steps)
    t_adj_stored = time.perf_counter() - t0
    stored_diff = float(np.abs(gz1 - gz).max() / np.abs(gz).max())
    # the mixed reverse sweep (float64 mode only): float32 pair terms for
    # well-separated tiles, recomputation with shared float64 reciprocals;
    # the forward sweep, hence J and the checkpoints, stays float64
    mixed = None
    if mode == "fp64":
        ratio = 0.7
        ctx.set_adjoint_precision(_lib.BVE_PRECISION_MIXED, ratio)
        gradient(ctx, F, g.active_area, dt, 2, 1)
        t0 = time.perf_counter()
        Jm, gzm, gxm = gradient(ctx, F, g.active_area, dt, steps, every)
        t_mixed = time.perf_counter() - t0
        ctx.set_adjoint_precision(_lib.BVE_PRECISION_FP64)
        mixed = dict(ratio=ratio, gradient_ms=t_mixed * 1e3,
                     adjoint_fwd_ratio_same_map=t_mixed / t_fwd_active,
                     J_equals_float64=bool(Jm == J),
                     grad_zeta_max_rel_diff_vs_float64=float(np.abs(gzm - gz).max() / np.abs(gz).max()),
                     grad_x_max_rel_diff_vs_float64=float(np.abs(gxm - gx).max() / np.abs(gx).max()))

    out = dict(workload=f"rh4_cubed_sphere_L{level}_{steps}steps_dt{dt:g}_checkpoint{every}",
               mode=mode, n_active=g.n_active, J=J, gradient_ms=t_adj * 1e3,
               forward_ms_active_only=t_fwd_active * 1e3,
               forward_ms_active_and_vertices=t_fwd_vertices * 1e3,
               forward_ms_all_classes=t_fwd * 1e3,
               adjoint_fwd_ratio_same_map=t_adj / t_fwd_active,
               gradient_ms_all_stages_stored=t_adj_stored * 1e3,
               adjoint_fwd_ratio_same_map_all_stages_stored=t_adj_stored / t_fwd_active,
               all_stages_stored_vs_checkpointed_max_rel_diff=stored_diff,
               adjoint_fwd_ratio_active_and_vertices=t_adj / t_fwd_vertices,
               adjoint_fwd_ratio_all_classes=t_adj / t_fwd,
               fd_panels=int(n_fd))
    if mixed is not None:
        out["mixed_reverse_sweep"] = mixed
    if n_fd <= 0:
        ctx.close()
        return out
    idx = np.random.default_rng(0).choice(g.n_active, n_fd, replace=False)

    def Jof(zz):
        load(zz, particles=False)
        xT, zT = ctx.adjoint_forward(dt, steps, steps)
        return math.fsum((g.active_area * zT * F.weight(xT)).tolist())

    def report(fd, grad=gz):
        rel = np.abs(grad[idx] - fd) / np.abs(fd)
        return dict(max_rel_err=float(rel.max()), median_rel_err=float(np.median(rel)),
                    normwise_rel_err=float(np.abs(grad[idx] - fd).max() / np.abs(fd).max()))

    # central differences at eps = 1e-6 with each J difference propagated in
    # difference form (bve_difference_forward): no cancellation
    load(z, particles=False)
    t0 = time.perf_counter()
    fd_exact = np.array([central_difference(ctx, F, g.active_area, dt, steps, j, 1e-6)
                         for j in idx])
    t_exact = time.perf_counter() - t0
    fd_report = {"eps_1e-06_difference_form": dict(report(fd_exact),
                                                   seconds=t_exact)}
    if mixed is not None:
        mixed["fd_eps_1e-06_difference_form"] = report(fd_exact, gzm)
    # plain quotients of two forward runs, for comparison: at eps = 1e-6 the
    # rounding noise of J (~1e-14 relative) dominates them
    for eps in (1e-6, 1e-3):
        fd = np.array([(Jof(z + eps * (np.arange(g.n_active) == j)) -
                        Jof(z - eps * (np.arange(g.n_active) == j))) / (2 * eps) for j in idx])
        fd_report[f"eps_{eps:g}_plain"] = report(fd)
    ctx.close()
    out["fd"] = fd_report
    return out


def ensemble_measurement(n_members=64, levels=(4, 5, 6), steps=10, dt=0.01,
                         modes=("mixed", "fp64_fast")):
    """Ensemble workload: 64 members cycling through cubed-sphere levels 4, 5
    and 6 (RH4 plus a seeded smooth perturbation of RMS 0.05 max|zeta|,
    correlation 0.3 rad), 10 RK4 steps at dt = 0.01 in one HIP graph, launch
    targets per lane tuned on the ensemble itself, segment lengths derived
    from its whole work.  Aggregate pair rate over all members (all three
    target classes counted)."""
    from thesiscode_amd import ensemble as ensmod
    members = ensmod.make_members(n_members, levels)
    out = {}
    for mode in modes:
        prec = {"mixed": 1, "fp64": 0, "fp64_fast": 2}[mode]
        table = ensmod.autotune_members(members, precision=prec)
        el, pairs, ens = ensmod.run(members, dt, steps, prec, table=table, use_graph=True)
        plans = {}
        for m, d in enumerate(members):
            plans.setdefault(str(d["level"]), list(ens.member_plan(m)))
        ens.close()
        out[mode] = dict(pair_rate_all_classes=pairs * steps / el, seconds=el,
                         ms_per_step=el * 1e3 / steps,
                         plans_targets_per_lane_chunk_work_items=plans,
                         items_per_workgroup=next(iter(table.values()))["items_per_workgroup"],
                         tuning_ms_per_stage={f"T{T}_items{k}": ms for T, k, ms in
                                              next(iter(table.values()))["tried"]})
    return dict(workload=f"{n_members} members, levels {list(levels)}, {steps} RK4 steps dt={dt}, "
                         f"one HIP graph", modes=out)


def class_sample(x, centres, n=64, seed=7):
    """Indices of the sampled targets of one class: the first n, the n
    nearest to each vortex centre (the steepest part of the field) and n
    seeded random ones."""
    idx = [np.arange(min(n, len(x)))]
    for c in centres:
        idx.append(np.argsort(1.0 - x @ c)[:n])
    idx.append(np.random.default_rng(seed).choice(len(x), min(n, len(x)), replace=False))
    return np.unique(np.concatenate(idx))


class ErrorSample:
    """Sampled targets of all three classes and their oracle velocities
    (float64, the reference's rounding order: BVEActiveRHS with
    self-exclusion by index, BVESmoothVelocity, BVEPassiveRHS)."""

    def __init__(self, grid, za, centres, smooth=0.05, n=64):
        import oracle
        x, a = grid.active_x, grid.active_area
        self.ia = class_sample(x, centres, n)
        self.ip = class_sample(grid.passive_x, centres, n)
        self.iq = class_sample(grid.particles_x, centres, n)
        self.ua = np.array([oracle.active_rhs(x, za, a, 0.0, i, i + 1)[0][i] for i in self.ia])
        self.up = oracle.smooth_velocity(grid.passive_x[self.ip], x, za, a, smooth)
        self.uq = oracle.passive_rhs(grid.particles_x[self.iq], x, za, a)

    def __len__(self):
        return len(self.ia) + len(self.ip) + len(self.iq)

    def errors(self, ctx):
        """Velocities of the resident state evaluated by the context's own
        pass (the plan and kernel a step times, over all targets at once)
        against the oracle at the sample.  Returns (normwise, max pointwise,
        per class, rms): normwise max_i |du_i| / max_i |u_i| over the whole
        sample, pointwise max_i |du_i| / |u_i| over targets with |u_i| >
        1e-3 max |u|, rms sqrt(mean |du_i|^2 / mean |u_i|^2)."""
        ua, up, uq = ctx.resident_velocity()
        got = np.concatenate([ua[self.ia], up[self.ip], uq[self.iq]])
        ref = np.concatenate([self.ua, self.up, self.uq])
        du = np.linalg.norm(got - ref, axis=1)
        un = np.linalg.norm(ref, axis=1)
        keep = un > 1e-3 * un.max()
        k1, k2 = len(self.ia), len(self.ia) + len(self.ip)
        per = {cls: float(du[sl].max() / un.max())
               for cls, sl in (("active", slice(0, k1)), ("passive", slice(k1, k2)),
                               ("particles", slice(k2, None)))}
        rms = float(np.sqrt(np.mean(du ** 2) / np.mean(un ** 2)))
        return float(du.max() / un.max()), float((du[keep] / un[keep]).max()), per, rms

    @staticmethod
    def positions(ctx):
        ax, _, px = ctx.store_panels()
        qx, _ = ctx.store_particles()
        return np.concatenate([ax, px, qx])

    def set_step_reference(self, ctx, reload, dt):
        """One float64 RK4 step of every target from the initial state (the
        context must be in the float64 mode): the reference of step_error,
        itself within ~1e-12 of the oracle's BVERK4."""
        reload()
        self.x0 = self.positions(ctx)
        ctx.advance(dt, 1)
        self.x1 = self.positions(ctx)
        self.disp = float(np.abs(self.x1 - self.x0).max())

    def step_error(self, ctx, reload, dt):
        """One RK4 step of the context's mode from the initial state (the
        advance path a timed run takes, all four stages) against the float64
        step: max position difference over every target / max displacement."""
        reload()
        ctx.advance(dt, 1)
        return float(np.abs(self.positions(ctx) - self.x1).max() / self.disp)


def peak_of(mode):
    """Vector peak of the pipe a mode's pair loop runs on (TFLOP/s)."""
    return PEAK_FP32_TFLOPS if mode.startswith("mixed") else PEAK_FP64_TFLOPS


def measure(ctx, reload, mode, args, pairs_primary, pairs_all, sample):
    """Times args.steps RK4 steps of one arithmetic mode from the initial
    state (reload() puts it back), then measures the mode's error at the
    initial state with the same plan."""
    from thesiscode_amd import _lib
    ctx.set_precision(MODES[mode])
    reload()
    ctx.advance(args.dt, args.warmup)
    ctx.synchronize()
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    ctx.advance(args.dt, args.steps)
    ctx.synchronize()
    el = time.perf_counter() - t0
    launches, kernel_ms = ctx.profile()
    ctx.set_profiling(False)
    T, tile, chunk, n_chunks, recip = ctx.plan()
    launch = ctx.launch()
    # HBM bytes per launch, modelled: the partial-sum rows written (24 B per
    # target per row), the targets each row's work item reads (the same
    # count) and the source records (32 B each); for the float64 segmented
    # launch at level 7 this is 206 MB against 199 MB measured.  The mixed
    # and bounded passes also fetch their source records again from HBM
    # (6.3 MB of them per pass exceed an XCD's L2), which the model leaves out.
    traffic_model = 2.0 * launch["partial_bytes"] + 32.0 * ctx.n_active
    kernel = kernel_name(mode, T, tile, recip, launch["segments"], ctx.n_active)
    # accuracy of this mode vs the oracle: the resident state's velocities
    # from the same plan and kernel as the timed launches
    reload()
    err, err_point, err_class, err_rms = sample.errors(ctx)
    step_err = sample.step_error(ctx, reload, args.dt)
    ctx.set_precision(_lib.BVE_PRECISION_FP64)
    avg_ms = kernel_ms / max(launches, 1)
    achieved = pairs_all / 4.0 * FLOPS_PER_PAIR / (avg_ms * 1e-3) / 1e12
    return dict(pair_rate=pairs_primary * args.steps / el,
                pair_rate_all_classes=pairs_all * args.steps / el,
                ms_per_step=el * 1e3 / max(args.steps, 1), rel_err_vs_fp64_oracle=err,
                max_pointwise_rel_err=err_point, rms_rel_err=err_rms, rel_err_per_class=err_class,
                step_rel_err_vs_fp64_step=step_err,
                error_sample=f"{len(sample)} targets (active, passive and particle classes), "
                             f"velocities of the resident state from the timed plan",
                launches=launches, avg_launch_ms=avg_ms, achieved_tflops=achieved,
                peak_tflops=peak_of(mode), roofline_frac=achieved / peak_of(mode),
                plan=dict(targets_per_lane=T, tile=tile, chunk=chunk, n_chunks=n_chunks,
                          recip=recip, **launch),
                kernel=kernel,
                traffic_bytes_per_launch=MEASURED_TRAFFIC.get((kernel, args.level)),
                traffic_model_bytes_per_launch=traffic_model)


def pareto_curve(ctx, reload, sample, dt, steps, base_launch_ms, base_step_ms):
    """Error against speed of the reduced-precision modes on one problem:
    the float32 mixed pass at near-field radii 0.02 to 0.5, with and
    without Kahan sums, and the bounded pass at error targets 1e-8 to
    1e-11.  Speed-ups are against the plain float64 pass (per launch and
    per RK4 step); errors are normwise (and max pointwise) over the
    sample's three classes, from the timed plan."""
    from thesiscode_amd import _lib
    points = [("mixed", r) for r in (0.02, 0.05, 0.1, 0.2, 0.5)]
    points += [("mixed_kahan", r) for r in (0.02, 0.1, 0.5)]
    points += [("mixed_bounded", t) for t in (1e-8, 1e-9, 1e-10, 1e-11)]
    out = []
    for mode, param in points:
        if mode == "mixed_bounded":
            ctx.set_precision(MODES[mode])
            ctx.set_error_target(param)
        else:
            ctx.set_precision(MODES[mode], param)
        reload()
        ctx.advance(dt, 1)
        ctx.synchronize()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        ctx.advance(dt, steps)
        ctx.synchronize()
        el = time.perf_counter() - t0
        launches, kernel_ms = ctx.profile()
        ctx.set_profiling(False)
        reload()
        err, err_point, _, err_rms = sample.errors(ctx)
        step_err = sample.step_error(ctx, reload, dt)
        launch_ms = kernel_ms / max(launches, 1)
        step_ms = el * 1e3 / steps
        out.append(dict(mode=mode, **({"error_target": param} if mode == "mixed_bounded"
                                      else {"nearfield_radius": param}),
                        ms_per_launch=launch_ms, ms_per_step=step_ms,
                        speedup_vs_fp64_launch=base_launch_ms / launch_ms,
                        speedup_vs_fp64_step=base_step_ms / step_ms,
                        rel_err=err, max_pointwise_rel_err=err_point, rms_rel_err=err_rms,
                        step_rel_err=step_err))
        print(f"# pareto {out[-1]}", file=sys.stderr, flush=True)
    ctx.set_error_target(1e-10)
    ctx.set_precision(_lib.BVE_PRECISION_FP64)
    ok = [p for p in out if max(p["rel_err"], p["step_rel_err"]) <= ERR_BAR]
    best = max(ok, key=lambda p: p["speedup_vs_fp64_step"]) if ok else None
    return {"points": out, "fastest_within_1e-10": best}


def host_roundtrip_measurement(ctx, grid, absv, dt, steps, pairs_primary, mode="fp64_fast"):
    """The drop-in's per-step pattern (the Fortran BVERK4): particles and
    their absolute vorticity uploaded, one RK4 step, particles, their
    relative vorticity and KE downloaded (bve_rk4_step), then the panels
    downloaded for ScatterPanels (bve_store_panels).  Host arrays are
    ordinary (pageable) memory, as Fortran's are."""
    from thesiscode_amd import _lib
    ctx.set_precision(MODES[mode])
    qx = np.ascontiguousarray(grid.particles_x).copy()
    qa = np.ascontiguousarray(absv)
    rel, ke = np.zeros(grid.n_particles), np.zeros(grid.n_active)
    ax, av = np.zeros((grid.n_active, 3)), np.zeros(grid.n_active)
    px = np.zeros((grid.n_passive, 3))
    ctx.rk4_step_into(dt, qx, qa, rel, ke)
    ctx.store_panels_into(ax, av, px)
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.rk4_step_into(dt, qx, qa, rel, ke)
        ctx.store_panels_into(ax, av, px)
    el = (time.perf_counter() - t0) / steps
    ctx.set_precision(_lib.BVE_PRECISION_FP64)
    up = 32.0 * grid.n_particles
    down = 32.0 * grid.n_particles + 32.0 * grid.n_active + 24.0 * grid.n_passive + 8.0 * grid.n_active
    return dict(mode=mode, ms_per_step=el * 1e3, pair_rate_pcie_inclusive=pairs_primary / el,
                host_bytes_up_per_step=up, host_bytes_down_per_step=down, steps=steps)


def large_solve_measurement(level=8, modes=("fp64", "fp64_fast", "mixed", "mixed_kahan", "mixed_bounded"),
                            dt=0.002):
    """Large single solve: two co-rotating Gaussian vortices (amplitude 4 PI,
    width 0.1, at 20N +-10E, minus the mean) on the level-8 cubed sphere
    (393,216 active panels, 393,218 vertices, 131,070 parents); one RK4 step
    at dt = 0.002 per mode after one warm-up step, state resident in HBM."""
    import types
    from thesiscode_amd import _lib
    from thesiscode_amd import grid as gridmod
    from thesiscode_amd import initial_conditions as ic
    g = gridmod.cubed_sphere(level)
    za, zp, _ = ic.two_vortices(g)
    omega = ic.OMEGA_EARTH
    absv = zp + 2 * omega * g.particles_x[:, 2]
    na, npart, npp = g.n_active, g.n_particles, g.n_passive
    pairs_primary = 4.0 * (na * (na - 1.0) + npart * na)
    pairs_all = 4.0 * (na * (na - 1.0) + (npart + npp) * na)
    lat, lon = np.deg2rad(20.0), np.deg2rad(10.0)
    sample = ErrorSample(g, za, [ic.latlon_to_xyz(lat, lon), ic.latlon_to_xyz(lat, -lon)], n=48)
    ctx = _lib.Context(0)
    ctx.set_omega(omega)

    def reload():
        ctx.load_panels(g.active_x, za, g.active_area, g.passive_x)
        ctx.load_particles(g.particles_x, absv)

    args = types.SimpleNamespace(dt=dt, warmup=1, steps=1, level=level)
    ctx.set_precision(MODES["fp64"])
    sample.set_step_reference(ctx, reload, dt)
    out = {}
    for mode in modes:
        out[mode] = measure(ctx, reload, mode, args, pairs_primary, pairs_all, sample)
        print(f"# large {mode}: {out[mode]}", file=sys.stderr, flush=True)
    curve = None
    if "fp64" in out:
        curve = pareto_curve(ctx, reload, sample, dt, 1, out["fp64"]["avg_launch_ms"],
                             out["fp64"]["ms_per_step"])
    ctx.close()
    return dict(workload=f"two_vortices_cubed_sphere_L{level}_rk4_step_dt{dt:g}", level=level,
                n_active=na, n_particles=npart, n_passive=npp, pairs_per_step=pairs_primary,
                pairs_per_step_all_classes=pairs_all, modes=out, mixed_pareto=curve)


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

    def reload():
        ctx.load_panels(g.active_x, za, g.active_area, g.passive_x)
        ctx.load_particles(g.particles_x, absv)

    reload()
    print(f"# level {args.level}: {na} active, {npart} particles, {npp} passive", file=sys.stderr,
          flush=True)

    sample = ErrorSample(g, za, [ic.latlon_to_xyz(np.pi / 4, 0.0)])
    ctx.set_precision(MODES["fp64"])
    sample.set_step_reference(ctx, reload, args.dt)
    results = {}
    for mode in [m for m in args.modes.split(",") if m]:
        results[mode] = measure(ctx, reload, mode, args, pairs_primary, pairs_all, sample)
        print(f"# {mode}: {results[mode]}", file=sys.stderr, flush=True)
    if "fp64" in results:
        curve = pareto_curve(ctx, reload, sample, args.dt, max(2, min(args.steps, 5)),
                             results["fp64"]["avg_launch_ms"], results["fp64"]["ms_per_step"])
    else:
        curve = None
    roundtrip = host_roundtrip_measurement(ctx, g, absv, args.dt, max(args.steps // 2, 1),
                                           pairs_primary)
    print(f"# host round trip: {roundtrip}", file=sys.stderr, flush=True)
    ctx.close()

    ok = [m for m in results if max(results[m]["rel_err_vs_fp64_oracle"],
                                    results[m]["step_rel_err_vs_fp64_step"]) <= ERR_BAR]
    best = max(ok, key=lambda m: results[m]["pair_rate"]) if ok else None
    r = results[best] if best else {}
    out = {
        "metric": METRIC,
        "value": r.get("pair_rate"),
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r.get("ms_per_step"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        # the arithmetic of the headline mode: float64, or float32 pair terms
        # with float64 expansions, near field and sums
        "dtype": "f32+f64" if (best or "").startswith("mixed") else "f64",
        "data": "synthetic",
        "config": {
            "workload": f"gauss_vortex_cubed_sphere_L{args.level}_rk4_step",
            "level": args.level, "n_active": na, "n_particles": npart, "n_passive": npp,
            "dt": args.dt, "mode": best,
            "pairs_per_step": pairs_primary, "pairs_per_step_all_classes": pairs_all,
            "modes": results,
            "mixed_pareto": curve,
            "host_roundtrip": roundtrip,
        },
        "roofline": {
            # compute-bound on the vector ALU of the mode's precision (the
            # float64 pair loop keeps the FP64 VALU busy 94-97% of cycles);
            # the pair sum is not a contraction and issues no MFMA instruction
            "bound": "vector_fp32" if (best or "").startswith("mixed") else "vector_fp64",
            "pipe": ("vector FP32 VALU" if (best or "").startswith("mixed") else "vector FP64 VALU")
                    + " (no MFMA instructions)",
            "achieved": r.get("achieved_tflops"),
            "peak": r.get("peak_tflops", PEAK_FP64_TFLOPS),
            "unit": "TFLOP/s",
            "frac": r.get("roofline_frac"),
            "traffic": r.get("traffic_bytes_per_launch"),
            "traffic_model": r.get("traffic_model_bytes_per_launch"),
            "traffic_source": "rocprofv3 FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00 per launch of this "
                              "kernel on this workload (copy-calibrated; null when not measured); "
                              "traffic_model: partial-sum rows written + the targets their "
                              "segments read + the source records",
            "kernel": r.get("kernel"),
            "launches": r.get("launches"),
            "avg_launch_ms": r.get("avg_launch_ms"),
            "flops_per_launch": pairs_all / 4.0 * FLOPS_PER_PAIR,
        },
    }
    if args.adjoint_level > 0:
        adj = adjoint_measurement(args.adjoint_level, n_fd=args.adjoint_fd_panels)
        print(f"# adjoint: {adj}", file=sys.stderr, flush=True)
        # the gradient against a forward of the map it differentiates (the
        # active panels alone, the same 20 steps and arithmetic)
        out["adjoint_fwd_ratio"] = adj["adjoint_fwd_ratio_same_map"]
        # the mixed reverse sweep's gradient (float64 forward sweep) against
        # the same float64 forward
        out["adjoint_fwd_ratio_mixed_reverse_sweep"] = adj["mixed_reverse_sweep"]["adjoint_fwd_ratio_same_map"]
        # the same in the shared-reciprocal float64 mode (its forward passes
        # and recomputation on pair_sum_f64_seg_kernel<..,4>)
        adj["fp64_fast"] = adjoint_measurement(args.adjoint_level, n_fd=0, mode="fp64_fast")
        print(f"# adjoint fp64_fast: {adj['fp64_fast']}", file=sys.stderr, flush=True)
        out["config"]["adjoint"] = adj
    if args.ensemble_members > 0:
        ens = ensemble_measurement(args.ensemble_members,
                                   modes=[m for m in args.ensemble_modes.split(",") if m])
        print(f"# ensemble: {ens}", file=sys.stderr, flush=True)
        out["config"]["ensemble"] = ens
    if args.large_level > 0:
        big = large_solve_measurement(args.large_level,
                                      modes=[m for m in args.large_modes.split(",") if m])
        out["config"]["large_solve"] = big
    if args.cpu_seconds > 0:
        port = cpu_baseline(g, za, args.cpu_seconds)
        # the headline size (level 7, one step: ~20 s on 16 host cores), and
        # three level-6 steps beside it
        ref = cpu_baseline_reference(level=args.level, steps=1)
        out["cpu_baseline"] = ref if ref is not None else port
        out["config"]["cpu_port"] = port
        out["config"]["cpu_reference_level6"] = cpu_baseline_reference(level=6, steps=3)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
