"""BVE_PRECISION_MIXED_BOUNDED: float64 second-order tile expansions plus
float32 remainders, with every (target group, tile) pair whose remainder
rounding bound exceeds the error target routed to float64.

* Error target 0: every pair takes the float64 path, so the result meets
  the float64 bar (pins the bookkeeping: no pair lost or counted twice,
  self pairs excluded, the smoothed class honoured, the probe and the
  routing harmless).
* The benchmark workloads (level-7 Gaussian vortex, level-8 two vortices):
  normwise error <= 1e-10 against the oracle on sampled targets of all
  three classes, on the default plan, and through one RK4 step.
* The expansion itself: with the near and heavy tiles still routed but a
  loose target, the error grows (the float32 remainders are really used)
  yet stays far below a plain float32 sum's ~4e-7."""
import numpy as np
import pytest

import oracle
from conftest import normwise_rel
from thesiscode_amd import _lib
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic

pytestmark = pytest.mark.gpu
BOUNDED = _lib.BVE_PRECISION_MIXED_BOUNDED
OMEGA = ic.OMEGA_EARTH
TARGET = 1e-10


def class_sample(x, centres, n=64, seed=7):
    idx = [np.arange(min(n, len(x)))]
    for c in centres:
        idx.append(np.argsort(1.0 - x @ c)[:n])
    idx.append(np.random.default_rng(seed).choice(len(x), min(n, len(x)), replace=False))
    return np.unique(np.concatenate(idx))


def sampled_errors(g, za, u, centres, n=64):
    """Normwise error of (u_active, u_passive, u_particles) against the
    oracle over sampled targets of all classes together."""
    x, a = g.active_x, g.active_area
    ia = class_sample(x, centres, n)
    ip = class_sample(g.passive_x, centres, n)
    iq = class_sample(g.particles_x, centres, n)
    ra = np.array([oracle.active_rhs(x, za, a, OMEGA, i, i + 1)[0][i] for i in ia])
    rp = oracle.smooth_velocity(g.passive_x[ip], x, za, a, 0.05)
    rq = oracle.passive_rhs(g.particles_x[iq], x, za, a)
    got = np.concatenate([u[0][ia], u[1][ip], u[2][iq]])
    ref = np.concatenate([ra, rp, rq])
    return normwise_rel(got, ref)


def bounded_context(g, za, absv, target=TARGET):
    c = _lib.Context(0)
    c.set_omega(OMEGA)
    c.set_precision(BOUNDED)
    c.set_error_target(target)
    c.load_panels(g.active_x, za, g.active_area, g.passive_x)
    c.load_particles(g.particles_x, absv)
    return c


def test_target_zero_is_float64():
    g = gridmod.cubed_sphere(4)
    z = ic.rh4_wave(g.active_x)
    na, npp = g.n_active, g.n_passive
    ref_a, _ = oracle.active_rhs(g.active_x, z, g.active_area, 0.0)
    ref_s = oracle.smooth_velocity(g.passive_x, g.active_x, z, g.active_area, 0.05)
    ref_p = oracle.passive_rhs(g.particles_x, g.active_x, z, g.active_area)
    t = np.concatenate([g.active_x, g.passive_x, g.particles_x])
    with _lib.Context(0) as c:
        c.set_precision(BOUNDED)
        c.set_error_target(0.0)
        u = c.velocity(t, g.active_x, z, g.active_area, self_end=na, smooth_range=(na, na + npp))
    assert normwise_rel(u[:na], ref_a) < 1e-12
    assert normwise_rel(u[na:na + npp], ref_s) < 1e-12
    assert normwise_rel(u[na + npp:], ref_p) < 1e-12


@pytest.mark.parametrize("chunk", [0, 1024])
def test_level7_gauss_within_target(chunk):
    """config[1]: level-7 Gaussian vortex, resident state on the default
    plan and on one with explicit 1024-source segments."""
    g = gridmod.cubed_sphere(7)
    za, zp, _ = ic.gauss_vortex(g, lat0=np.pi / 4, lon0=0.0, width=0.05)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    with bounded_context(g, za, absv) as c:
        c.set_chunk(chunk)
        assert c.launch()["segments"] > 0
        # two targets per lane: pair_sum_bounded_seg_kernel<2,true>, the kernel
        # bench.py times
        assert c.plan()[0] == 2
        if chunk:
            assert c.plan()[2] == chunk
        u = c.resident_velocity()
    err = sampled_errors(g, za, u, [ic.latlon_to_xyz(np.pi / 4, 0.0)])
    print(f"level 7 bounded (chunk {chunk}): {err:.3e}")
    assert err < TARGET


@pytest.mark.parametrize("target", [1e-8, 1e-9, 1e-11])
def test_level6_within_each_target(target):
    """The contract at other targets: every routing regime (whole tiles in
    float32, sub-tiles retried in float32 or float64, float64 tiles) mixes
    differently with the target, and the error stays within it.  Level 6
    runs one target per lane."""
    g = gridmod.cubed_sphere(6)
    za, zp, _ = ic.gauss_vortex(g, lat0=np.pi / 4, lon0=0.0, width=0.05)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    with bounded_context(g, za, absv, target) as c:
        assert c.plan()[0] == 1
        u = c.resident_velocity()
    err = sampled_errors(g, za, u, [ic.latlon_to_xyz(np.pi / 4, 0.0)])
    print(f"level 6 bounded, target {target:g}: {err:.3e}")
    assert err < target


def test_level8_two_vortices_within_target():
    """config[4]: level-8 two vortices."""
    g = gridmod.cubed_sphere(8)
    za, zp, _ = ic.two_vortices(g)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    lat, lon = np.deg2rad(20.0), np.deg2rad(10.0)
    with bounded_context(g, za, absv) as c:
        u = c.resident_velocity()
    err = sampled_errors(g, za, u, [ic.latlon_to_xyz(lat, lon), ic.latlon_to_xyz(lat, -lon)],
                         n=48)
    print(f"level 8 bounded: {err:.3e}")
    assert err < TARGET


@pytest.mark.parametrize("T", [1, 4])
def test_level8_staged_records_every_plan(T):
    """Above 131,072 sources the bounded pass stages its float32 records in
    LDS instead of reading them through scalar loads (a second kernel
    instantiation): the level-8 two-vortex velocities on the 1- and
    4-targets-per-lane plans (the default test covers 2) meet the target
    against the oracle too."""
    g = gridmod.cubed_sphere(8)
    za, zp, _ = ic.two_vortices(g)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    lat, lon = np.deg2rad(20.0), np.deg2rad(10.0)
    with bounded_context(g, za, absv) as c:
        c.set_plan(T)
        assert c.plan()[0] == T
        u = c.resident_velocity()
    err = sampled_errors(g, za, u, [ic.latlon_to_xyz(lat, lon), ic.latlon_to_xyz(lat, -lon)],
                         n=32)
    print(f"level 8 bounded, {T} targets per lane: {err:.3e}")
    assert err < TARGET


def test_level8_step_within_target():
    """config[4]: one RK4 step of the level-8 two-vortex state at dt = 0.002
    in the bounded mode against the float64 step (itself held to the oracle
    on sampled targets in test_gpu_workloads): every target's position
    within 1e-10 of the step's displacement."""
    g = gridmod.cubed_sphere(8)
    za, zp, _ = ic.two_vortices(g)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    x0 = np.concatenate([g.active_x, g.passive_x, g.particles_x])
    out = {}
    for precision in (_lib.BVE_PRECISION_FP64, BOUNDED):
        print("level 8 step, precision", precision, flush=True)
        with bounded_context(g, za, absv) as c:
            c.set_precision(precision)
            c.advance(0.002, 1)
            ax, _, px = c.store_panels()
            qx, _ = c.store_particles()
        out[precision] = np.concatenate([ax, px, qx])
    disp = np.abs(out[_lib.BVE_PRECISION_FP64] - x0).max()
    err = np.abs(out[BOUNDED] - out[_lib.BVE_PRECISION_FP64]).max() / disp
    print(f"level 8 bounded step: {err:.3e}")
    assert err < TARGET


def test_level7_step_within_target():
    """One RK4 step at level 7 in the bounded mode against the oracle's
    BVERK4 (every active panel, sampled parents and vertices)."""
    g = gridmod.cubed_sphere(7)
    za, zp, _ = ic.gauss_vortex(g, lat0=np.pi / 4, lon0=0.0, width=0.05)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    c0 = ic.latlon_to_xyz(np.pi / 4, 0.0)
    ip, iq = class_sample(g.passive_x, [c0]), class_sample(g.particles_x, [c0])
    dt = 0.005
    with bounded_context(g, za, absv) as c:
        c.advance(dt, 1)
        qx, _ = c.store_particles()
        ax, av, px = c.store_panels()
    r = oracle.rk4_step(g.active_x, za, g.active_area, g.passive_x[ip], g.particles_x[iq],
                        absv[iq], dt, OMEGA, 0.05)
    disp = np.linalg.norm(r["active_x"] - g.active_x, axis=1).max()
    err = max(np.abs(ax - r["active_x"]).max(), np.abs(px[ip] - r["passive_x"]).max(),
              np.abs(qx[iq] - r["particles_x"]).max()) / disp
    print(f"level 7 bounded step: {err:.3e}")
    assert err < TARGET


def test_loose_target_uses_the_remainders():
    """A loose target keeps more tiles in float32: the error rises above the
    float64 level but stays far below plain float32 terms."""
    g = gridmod.cubed_sphere(6)
    za, zp, _ = ic.gauss_vortex(g, width=0.05)
    absv = zp + 2 * OMEGA * g.particles_x[:, 2]
    errs = []
    for target in (1e-13, 1e-6):
        with bounded_context(g, za, absv, target) as c:
            u = c.resident_velocity()
        errs.append(sampled_errors(g, za, u, [ic.latlon_to_xyz(np.pi / 4, 0.0)]))
    print("bounded level 6, targets 1e-13 / 1e-6:", errs)
    assert errs[0] < 1e-12
    assert errs[1] < 1e-7


@pytest.mark.parametrize("na,npp,npart", [(1, 0, 0), (7, 3, 5), (300, 0, 257), (1029, 511, 1537)])
def test_ragged_inputs(na, npp, npart):
    """Sizes that fill no tile, segment or group evenly, random points: one
    short RK4 step of all classes against the oracle, to the float64 bar
    from the closest pair's conditioning or the error target."""
    rng = np.random.default_rng(na + npp + npart)

    def sphere(n, seed):
        p = np.random.default_rng(seed).normal(size=(n, 3))
        return p / np.linalg.norm(p, axis=1)[:, None]
    ax, px, qx = sphere(na, 1), sphere(npp, 2), sphere(npart, 3)
    zv, area = rng.normal(size=na), rng.uniform(0.5, 1.5, na) * 4 * np.pi / max(na, 1)
    qa = rng.normal(size=npart)
    with _lib.Context(0) as c:
        c.set_precision(BOUNDED)
        c.load_panels(ax, zv, area, px)
        qx1, _, _, _ = c.rk4_step(1e-5, qx, qa)
        ax1, _, px1 = c.store_panels()
    r = oracle.rk4_step(ax, zv, area, px, qx, qa, 1e-5, 2 * np.pi, 0.05)
    t = np.concatenate([ax, px, qx])
    d = 1.0 - t @ ax.T
    d[np.arange(na), np.arange(na)] = np.inf
    bound = max(TARGET, 16 * np.finfo(float).eps / d.min())
    for got, x0, key in ((ax1, ax, "active_x"), (px1, px, "passive_x"), (qx1, qx, "particles_x")):
        if len(got):
            disp = max(np.abs(r[key] - x0).max(), 1e-300)
            err = np.abs(got - r[key]).max() / disp
            print(f"{key}: {err:.2e} (bar {bound:.1e})")
            assert err < bound, key


def test_rh4_ten_steps_track_float64():
    """A smooth field (RH4 at level 6, all three target classes) over 10
    RK4 steps at dt = 0.01: the bounded mode's positions and vorticity stay
    within 10 x 1e-10 of the float64 run's, relative to their change over
    the run (the per-evaluation target, accumulated over the steps)."""
    g = gridmod.cubed_sphere(6)
    z = ic.rh4_wave(g.active_x)
    omega = ic.OMEGA_EARTH
    absv = ic.rh4_wave(g.particles_x) + 2 * omega * g.particles_x[:, 2]
    out = {}
    for precision in (_lib.BVE_PRECISION_FP64, BOUNDED):
        print("RH4 level 6, 10 steps, precision", precision, flush=True)
        with _lib.Context(0) as c:
            c.set_omega(omega)
            c.set_precision(precision)
            c.load_panels(g.active_x, z, g.active_area, g.passive_x)
            c.load_particles(g.particles_x, absv)
            c.advance(0.01, 10)
            ax, av, px = c.store_panels()
            qx, _ = c.store_particles()
        out[precision] = (np.concatenate([ax, px, qx]), av)
    (x64, z64), (xb, zb) = out[_lib.BVE_PRECISION_FP64], out[BOUNDED]
    disp = np.abs(x64 - np.concatenate([g.active_x, g.passive_x, g.particles_x])).max()
    pos = np.abs(xb - x64).max() / disp
    vort = np.abs(zb - z64).max() / np.abs(z64 - z).max()
    print(f"bounded RH4 L6, 10 steps: positions {pos:.2e}, vorticity {vort:.2e}")
    assert pos < 1e-9
    assert vort < 1e-9
