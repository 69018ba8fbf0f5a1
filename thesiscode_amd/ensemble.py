"""Ensembles of uneven panel problems on one GPU, and the launch-plan
autotuner they use.

The ensemble workload: member m lives on the cubed sphere of level
levels[m % len(levels)]; its relative vorticity is the RH4 wave plus a seeded
smooth random perturbation (sum of Gaussian bumps of width `corr` at random
centres, scaled to RMS sigma = 0.05 max|zeta| over the active panels).
"""
import json
import os
import time

import numpy as np

from . import _lib
from . import grid as gridmod
from . import initial_conditions as ic


def perturbation(grid, seed, sigma, corr=0.3, n_bumps=64):
    """Seeded smooth random field evaluated at the active panels, particles
    and passive panels of `grid`; one scaling (RMS sigma over the active
    panels) for all three."""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n_bumps, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    amps = rng.normal(size=n_bumps)

    def f(x):
        if len(x) == 0:
            return np.zeros(0)
        return np.exp(-2.0 * (1.0 - x @ c.T) / (corr * corr)) @ amps
    fa = f(grid.active_x)
    scale = sigma / np.sqrt(np.mean(fa * fa))
    return fa * scale, f(grid.particles_x) * scale, f(grid.passive_x) * scale


def make_members(n_members=64, levels=(4, 5, 6), seed=0, omega=ic.OMEGA_EARTH):
    """The ensemble's member inputs (dicts of arrays), member m at level
    levels[m % len(levels)]."""
    grids = {L: gridmod.cubed_sphere(L) for L in set(levels)}
    members = []
    for m in range(n_members):
        L = levels[m % len(levels)]
        g = grids[L]
        za = ic.rh4_wave(g.active_x)
        zp = ic.rh4_wave(g.particles_x)
        zq = ic.rh4_wave(g.passive_x) if g.n_passive else np.zeros(0)
        pa, pp, pq = perturbation(g, seed + m, 0.05 * np.abs(za).max())
        members.append(dict(level=L, grid=g, zeta_active=za + pa, zeta_particles=zp + pp,
                            zeta_passive=zq + pq,
                            absvort=zp + pp + 2 * omega * g.particles_x[:, 2]))
    return members


def load_members(ens, members):
    for m, d in enumerate(members):
        g = d["grid"]
        ens.load_member(m, g.active_x, d["zeta_active"], g.active_area, g.passive_x,
                        g.particles_x, d["absvort"])


def candidate_plans(level):
    """(targets per lane, chunk) pairs the autotuner tries for a level; chunk
    0 lets the ensemble derive the segment length from its whole work."""
    n_src = gridmod.counts(level)[0]
    chunks = [c for c in (256, 512, 1024, 2048, 4096, 8192, 16384) if c <= max(512, n_src)]
    return [(T, c) for T in (1, 2, 4) for c in [0] + chunks]


ITEMS_PER_WORKGROUP = (16, 32, 64, 128)


def autotune_members(members, precision=_lib.BVE_PRECISION_FP64, omega=ic.OMEGA_EARTH, dt=0.01,
                     device=0, log=None, steps=2):
    """Tune the ensemble itself: every member gets one targets-per-lane value
    (one shared pair launch) and chunk 0, the segment length the ensemble
    derives from its whole work and the work items per resident workgroup;
    each (T, items) in (1, 2, 4) x ITEMS_PER_WORKGROUP is timed over `steps`
    RK4 steps of all members from the initial state.  Returns {level:
    {"targets_per_lane", "chunk", "items_per_workgroup", "ms_per_stage",
    "tried"}} for the levels present (tried: (T, items, ms per stage))."""
    tried = []
    with _lib.Ensemble(len(members), device) as ens:
        ens.set_params(omega, 0.05, precision, 0.02)
        for T in (1, 2, 4):
            for items in ITEMS_PER_WORKGROUP:
                load_members(ens, members)
                ens.set_items_per_workgroup(items)
                for m in range(len(members)):
                    ens.set_member_plan(m, T, 0)
                ens.advance(dt, 1, use_graph=False)
                ens.synchronize()
                t0 = time.perf_counter()
                ens.advance(dt, steps, use_graph=False)
                ens.synchronize()
                ms = (time.perf_counter() - t0) * 1e3 / (4 * steps)
                tried.append((T, items, ms))
                if log:
                    log(f"autotune ensemble: T={T} items/workgroup={items} {ms:.3f} ms/stage")
    T, items, ms = min(tried, key=lambda r: r[2])
    return {L: dict(targets_per_lane=T, chunk=0, items_per_workgroup=items, ms_per_stage=ms,
                    tried=tried)
            for L in sorted({d["level"] for d in members})}


def autotune(levels, precision=_lib.BVE_PRECISION_FP64, members_per_level=2, omega=ic.OMEGA_EARTH,
             dt=0.01, device=0, log=None, weights=None):
    """Time every candidate plan on a homogeneous ensemble of each level
    (whole RK4 steps: the pair launches and the stage updates, which sum one
    partial row per chunk, so short chunks cost there what they save in
    balance) and return {level: {"targets_per_lane", "chunk",
    "ms_per_stage", "tried"}}.

    Members with the same targets per lane share one pair launch, so the
    table gives every level one common value: the T whose best chunks
    minimise sum_L weights[L] * ms_per_stage(L) (weights: members per level,
    default equal), and per level the fastest chunk at that T."""
    table = {}
    for L in levels:
        mem = make_members(members_per_level, (L,))
        tried = []
        with _lib.Ensemble(members_per_level, device) as ens:
            ens.set_params(omega, 0.05, precision, 0.02)
            load_members(ens, mem)
            for T, chunk in candidate_plans(L):
                # every candidate times the same steps from the initial state
                # (carried on through all candidates, ~80 steps, a perturbed
                # state rolls up and leaves the sphere)
                load_members(ens, mem)
                for m in range(members_per_level):
                    ens.set_member_plan(m, T, chunk)
                ens.advance(dt, 1, use_graph=False)
                ens.synchronize()
                steps = 3
                t0 = time.perf_counter()
                ens.advance(dt, steps, use_graph=False)
                ens.synchronize()
                per = (time.perf_counter() - t0) * 1e3 / (4 * steps)
                tried.append((T, chunk, per))
        table[L] = dict(tried=tried)
    for L, (T, chunk, ms) in common_plans({L: t["tried"] for L, t in table.items()},
                                          weights).items():
        table[L].update(targets_per_lane=T, chunk=chunk, ms_per_stage=ms)
        if log:
            log(f"autotune level {L}: T={T} chunk={chunk} {ms:.3f} ms/stage")
    return table


def common_plans(tried, weights=None):
    """From {level: [(T, chunk, ms per stage), ...]} pick the one T that
    minimises sum_L weights[L] * (fastest ms at that T on level L) among the
    values tried on every level, and per level its fastest (T, chunk, ms)."""
    weights = weights or {L: 1.0 for L in tried}

    def best_at(L, T):
        runs = [r for r in tried[L] if r[0] == T]
        return min(runs, key=lambda r: r[2]) if runs else None

    common = [T for T in (1, 2, 4) if all(best_at(L, T) for L in tried)]
    if not common:
        raise ValueError("no targets-per-lane value was tried on every level")
    T_best = min(common, key=lambda T: sum(weights.get(L, 1.0) * best_at(L, T)[2]
                                           for L in tried))
    return {L: best_at(L, T_best) for L in tried}


def save_table(table, path):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in table.items()}, f, indent=1)


def load_table(path):
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return {int(k): v for k, v in json.load(f).items()}


def run(members, dt, n_steps, precision, table=None, use_graph=True, omega=ic.OMEGA_EARTH,
        device=0, warmup=1):
    """Advance the ensemble n_steps after one untimed launch of the same
    n_steps (which also captures the graph when asked), and return
    (seconds, pairs per step, ensemble)."""
    ens = _lib.Ensemble(len(members), device)
    ens.set_params(omega, 0.05, precision, 0.02)
    load_members(ens, members)
    for m, d in enumerate(members):
        if table and d["level"] in table:
            t = table[d["level"]]
            if "items_per_workgroup" in t:
                ens.set_items_per_workgroup(t["items_per_workgroup"])
            ens.set_member_plan(m, t["targets_per_lane"], t["chunk"])
    if warmup:
        ens.advance(dt, n_steps, use_graph)
        ens.synchronize()
    t0 = time.perf_counter()
    ens.advance(dt, n_steps, use_graph)
    ens.synchronize()
    el = time.perf_counter() - t0
    return el, ens.pairs_per_step(), ens
