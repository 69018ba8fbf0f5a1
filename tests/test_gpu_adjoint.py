"""GPU adjoint: the reverse sweep against the float64 CPU tangent-linear
model (exact identity <grad, v> = dJ[v] for a linear functional), against
central finite differences of the GPU forward map (plain, and propagated in
difference form, which keeps epsilon = 1e-6 free of cancellation), and
checkpoint invariance."""
import numpy as np
import pytest

import oracle
from thesiscode_amd import _lib
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic
from thesiscode_amd.adjoint import GaussianWeightFunctional, central_difference, gradient

pytestmark = pytest.mark.gpu
OMEGA = 2 * np.pi
BAR_MIXED_TLM = 1e-10   # measured 3e-13 and 2.9e-12 at level 5


@pytest.fixture(scope="module")
def ctx():
    c = _lib.Context(0)
    c.set_omega(OMEGA)
    yield c
    c.close()


def load(ctx, level):
    g = gridmod.cubed_sphere(level)
    z = ic.rh4_wave(g.active_x)
    ctx.reset()
    ctx.load_panels(g.active_x, z, g.active_area, g.passive_x)
    return g, z


@pytest.mark.parametrize("precision", [_lib.BVE_PRECISION_FP64, _lib.BVE_PRECISION_FP64_FAST])
def test_adjoint_matches_tangent_linear(ctx, precision):
    """Linear functional J = <c_x, x_T> + <c_z, zeta_T>: the adjoint gradient
    contracted with a direction v equals the CPU tangent-linear dJ[v]."""
    g, z = load(ctx, 3)
    rng = np.random.default_rng(1)
    cx = rng.normal(size=(g.n_active, 3))
    cz = rng.normal(size=g.n_active)
    dt, nsteps = 0.01, 4
    ctx.set_precision(precision)
    try:
        xT, zT = ctx.adjoint_forward(dt, nsteps, 2)
        gx, gz = ctx.adjoint_backward(cx, cz)
    finally:
        ctx.set_precision(_lib.BVE_PRECISION_FP64)
    for trial in range(3):
        vx = rng.normal(size=(g.n_active, 3)) * 1e-2 * (trial > 0)
        vz = rng.normal(size=g.n_active)
        x1, z1, dxT, dzT = oracle.tlm_rk4(g.active_x, z, g.active_area, vx, vz, dt, OMEGA, nsteps)
        tlm = np.sum(cx * dxT) + np.sum(cz * dzT)
        adj = np.sum(gx * vx) + np.sum(gz * vz)
        print(f"trial {trial}: tlm {tlm:.15e} adj {adj:.15e}")
        assert abs(adj - tlm) <= 1e-10 * (abs(np.sum(np.abs(cx * dxT))) + abs(np.sum(np.abs(cz * dzT))))
    assert np.abs(xT - x1).max() < 1e-12


@pytest.mark.parametrize("eps,bound", [(1e-6, 1e-4), (1e-3, 1e-6)])
def test_adjoint_matches_finite_differences(ctx, eps, bound):
    """Gaussian-weight functional (width 0.2 at 30N 90E): dJ/dzeta_j(0) against
    central differences of the GPU forward map on 8 seeded panels.  The
    difference quotient carries rounding noise ~1e-16 |J| / eps, so at
    eps = 1e-6 it agrees to ~1e-5 and at eps = 1e-3 (truncation still
    negligible: J is nearly linear in one panel's vorticity) to < 1e-6."""
    import math
    g, z = load(ctx, 4)
    F = GaussianWeightFunctional()
    dt, nsteps, every = 0.01, 6, 2
    J, gz, gx = gradient(ctx, F, g.active_area, dt, nsteps, every)
    idx = np.random.default_rng(0).choice(g.n_active, 8, replace=False)
    for j in idx:
        vals = []
        for sgn in (1.0, -1.0):
            zz = z.copy()
            zz[j] += sgn * eps
            ctx.reset()
            ctx.load_panels(g.active_x, zz, g.active_area, g.passive_x)
            xT, zT = ctx.adjoint_forward(dt, nsteps, nsteps)
            vals.append(math.fsum((g.active_area * zT * F.weight(xT)).tolist()))
        fd = (vals[0] - vals[1]) / (2 * eps)
        rel = abs(gz[j] - fd) / abs(fd)
        print(f"eps {eps:g} panel {j}: adjoint {gz[j]:.12e} fd {fd:.12e} rel {rel:.2e}")
        assert rel < bound


def test_difference_forward_equals_plain_difference(ctx):
    """bve_difference_forward at a large perturbation (1e-2 of the vorticity
    on every panel), where two plain runs lose only ~1e-14 of the difference
    to cancellation: the propagated difference equals x_T(zeta + dzeta) -
    x_T(zeta) and the same for zeta, and its base trajectory is the forward
    run's."""
    g, z = load(ctx, 3)
    dt, nsteps = 0.01, 4
    dz = 1e-2 * np.random.default_rng(3).normal(size=g.n_active)
    xb, zb = ctx.adjoint_forward(dt, nsteps, nsteps)
    x0, z0, dxT, dzT = ctx.difference_forward(dt, nsteps, dz)
    ctx.reset()
    ctx.load_panels(g.active_x, z + dz, g.active_area, g.passive_x)
    xp, zp = ctx.adjoint_forward(dt, nsteps, nsteps)
    assert np.abs(x0 - xb).max() < 1e-13 and np.abs(z0 - zb).max() < 1e-12
    ex = np.abs(dxT - (xp - xb)).max() / np.abs(xp - xb).max()
    ez = np.abs(dzT - (zp - zb)).max() / np.abs(zp - zb).max()
    print(f"difference form vs two runs: x {ex:.2e}, zeta {ez:.2e}")
    assert ex < 1e-11 and ez < 1e-11


def test_difference_forward_tends_to_the_tangent_linear(ctx):
    """At a perturbation of 1e-9, where plain differences of two runs keep
    only ~7 digits, the propagated difference divided by the amplitude
    equals the CPU float64 tangent-linear model to O(amplitude)."""
    g, z = load(ctx, 3)
    dt, nsteps, amp = 0.01, 4, 1e-9
    v = np.random.default_rng(4).normal(size=g.n_active)
    _, _, dxT, dzT = ctx.difference_forward(dt, nsteps, amp * v)
    _, _, tx, tz = oracle.tlm_rk4(g.active_x, z, g.active_area, np.zeros((g.n_active, 3)), v, dt,
                                  OMEGA, nsteps)
    ex = np.abs(dxT / amp - tx).max() / np.abs(tx).max()
    ez = np.abs(dzT / amp - tz).max() / np.abs(tz).max()
    print(f"difference / amplitude vs tangent-linear: x {ex:.2e}, zeta {ez:.2e}")
    assert ex < 1e-6 and ez < 1e-6


def test_adjoint_matches_exact_central_differences(ctx):
    """The Gaussian-weight functional's dJ/dzeta_j(0) against central
    differences at epsilon = 1e-6 computed in difference form, on 8 seeded
    panels: agreement to 1e-6 per panel (the measured agreement is far
    closer; the difference quotient is now limited by its O(eps^2)
    truncation, not by cancellation)."""
    g, z = load(ctx, 4)
    F = GaussianWeightFunctional()
    dt, nsteps, every, eps = 0.01, 6, 2, 1e-6
    J, gz, gx = gradient(ctx, F, g.active_area, dt, nsteps, every)
    idx = np.random.default_rng(0).choice(g.n_active, 8, replace=False)
    for j in idx:
        fd = central_difference(ctx, F, g.active_area, dt, nsteps, j, eps)
        rel = abs(gz[j] - fd) / abs(fd)
        print(f"panel {j}: adjoint {gz[j]:.15e} central difference {fd:.15e} rel {rel:.2e}")
        assert rel < 1e-6


def mixed_gradient(ctx, F, area, dt, nsteps, every, ratio=0.7):
    ctx.set_adjoint_precision(_lib.BVE_PRECISION_MIXED, ratio)
    try:
        return gradient(ctx, F, area, dt, nsteps, every)
    finally:
        ctx.set_adjoint_precision(_lib.BVE_PRECISION_FP64)


@pytest.mark.parametrize("level,ratio", [(3, 0.7), (4, 0.7), (5, 0.7), (5, 0.9)])
def test_mixed_adjoint_matches_the_float64_gradient(ctx, level, ratio):
    """The mixed reverse sweep (float32 pair terms for well-separated tiles,
    shared-reciprocal recomputation) against the all-float64 gradient of the
    same forward sweep: the functional is the same bits, the gradient agrees
    to 1e-7 of its largest entry (measured 4e-9 at level 5).  Level 3
    (384 panels) pads its records to 1,024: whole padding tiles."""
    g, z = load(ctx, level)
    F = GaussianWeightFunctional()
    J64, gz64, gx64 = gradient(ctx, F, g.active_area, 0.01, 6, 2)
    J, gz, gx = mixed_gradient(ctx, F, g.active_area, 0.01, 6, 2, ratio)
    ez = np.abs(gz - gz64).max() / np.abs(gz64).max()
    ex = np.abs(gx - gx64).max() / np.abs(gx64).max()
    print(f"level {level} ratio {ratio}: zeta {ez:.2e}, x {ex:.2e}")
    assert J == J64
    assert ez < 1e-7 and ex < 1e-7


def test_mixed_adjoint_ragged_count(ctx):
    """The first 5,144 panels of level 5 (a spatially coherent set whose
    far tiles run in float32): the records pad to 6,144, so the last
    source tile is partly padding and three are all padding; the mixed
    gradient still agrees with the float64 one to 1e-7."""
    g = gridmod.cubed_sphere(5)
    n = 5144
    x, a = g.active_x[:n], g.active_area[:n]
    z = ic.rh4_wave(x)
    ctx.reset()
    ctx.load_panels(x, z, a, np.zeros((0, 3)))
    F = GaussianWeightFunctional()
    J64, gz64, gx64 = gradient(ctx, F, a, 0.01, 4, 2)
    J, gz, gx = mixed_gradient(ctx, F, a, 0.01, 4, 2)
    ez = np.abs(gz - gz64).max() / np.abs(gz64).max()
    ex = np.abs(gx - gx64).max() / np.abs(gx64).max()
    print(f"{n} panels: zeta {ez:.2e}, x {ex:.2e}")
    assert J == J64
    assert 0 < ez < 1e-7 and ex < 1e-7


def test_mixed_adjoint_matches_exact_central_differences(ctx):
    """The mixed gradient's dJ/dzeta_j(0) against central differences of the
    float64 forward map in difference form (eps = 1e-6) on 8 seeded panels,
    to 1e-6 per panel, the bar of the float64 gradient's test."""
    g, z = load(ctx, 5)
    F = GaussianWeightFunctional()
    dt, nsteps, every, eps = 0.01, 6, 2, 1e-6
    J, gz, gx = mixed_gradient(ctx, F, g.active_area, dt, nsteps, every)
    idx = np.random.default_rng(0).choice(g.n_active, 8, replace=False)
    for j in idx:
        fd = central_difference(ctx, F, g.active_area, dt, nsteps, j, eps)
        rel = abs(gz[j] - fd) / abs(fd)
        print(f"panel {j}: mixed adjoint {gz[j]:.15e} central difference {fd:.15e} rel {rel:.2e}")
        assert rel < 1e-6


def test_mixed_adjoint_is_deterministic(ctx):
    """Fixed-order sums throughout: two mixed gradients are the same bits,
    and so are two checkpoint intervals' (the recomputation is the same
    shared-reciprocal pass for every segment)."""
    g, z = load(ctx, 4)
    F = GaussianWeightFunctional()
    a = mixed_gradient(ctx, F, g.active_area, 0.01, 6, 2)
    b = mixed_gradient(ctx, F, g.active_area, 0.01, 6, 2)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def test_mixed_adjoint_matches_tangent_linear(ctx):
    """The mixed reverse sweep at level 5 (where far tiles run in float32)
    against the CPU float64 tangent-linear model, for a linear functional:
    <grad, v> = dJ[v] for random directions, to 1e-10 of the sum of the
    terms' magnitudes (the float64 sweep's bar)."""
    g, z = load(ctx, 5)
    rng = np.random.default_rng(5)
    cx = rng.normal(size=(g.n_active, 3))
    cz = rng.normal(size=g.n_active)
    dt, nsteps = 0.01, 4
    ctx.set_adjoint_precision(_lib.BVE_PRECISION_MIXED)
    try:
        ctx.adjoint_forward(dt, nsteps, 2)
        gx, gz = ctx.adjoint_backward(cx, cz)
    finally:
        ctx.set_adjoint_precision(_lib.BVE_PRECISION_FP64)
    worst = 0.0
    for trial in range(2):
        vx = rng.normal(size=(g.n_active, 3)) * 1e-2 * trial
        vz = rng.normal(size=g.n_active)
        _, _, dxT, dzT = oracle.tlm_rk4(g.active_x, z, g.active_area, vx, vz, dt, OMEGA, nsteps)
        tlm = np.sum(cx * dxT) + np.sum(cz * dzT)
        adj = np.sum(gx * vx) + np.sum(gz * vz)
        scale = np.sum(np.abs(cx * dxT)) + np.sum(np.abs(cz * dzT))
        worst = max(worst, abs(adj - tlm) / scale)
        print(f"trial {trial}: tlm {tlm:.15e} mixed adjoint {adj:.15e} rel {abs(adj - tlm) / scale:.2e}")
    assert worst < BAR_MIXED_TLM


def test_mixed_adjoint_repeated_reverse_sweep(ctx):
    """A second mixed reverse sweep after the same forward sweep recomputes
    the last checkpoint segment (the first used the forward sweep's own
    float64 stage inputs) with the shared-reciprocal pass: it agrees with
    the first to 1e-10 of the largest entry, not bit for bit."""
    g, z = load(ctx, 5)
    F = GaussianWeightFunctional()
    ctx.set_adjoint_precision(_lib.BVE_PRECISION_MIXED)
    try:
        xT, zT = ctx.adjoint_forward(0.01, 6, 2)
        lam_x, lam_zeta = F.cotangent(xT, zT, g.active_area)
        gx1, gz1 = ctx.adjoint_backward(lam_x, lam_zeta)
        gx2, gz2 = ctx.adjoint_backward(lam_x, lam_zeta)
    finally:
        ctx.set_adjoint_precision(_lib.BVE_PRECISION_FP64)
    ez = np.abs(gz2 - gz1).max() / np.abs(gz1).max()
    ex = np.abs(gx2 - gx1).max() / np.abs(gx1).max()
    print(f"repeated mixed reverse sweep: zeta {ez:.2e}, x {ex:.2e}")
    assert ez < 1e-10 and ex < 1e-10


def test_adjoint_precision_arguments(ctx):
    with pytest.raises(_lib.BVEError, match="precision must be"):
        ctx.set_adjoint_precision(_lib.BVE_PRECISION_FP64_FAST)
    for bad in (0.0, 1.0, -0.5, float("nan")):
        with pytest.raises(_lib.BVEError, match="ratio"):
            ctx.set_adjoint_precision(_lib.BVE_PRECISION_MIXED, bad)
    ctx.set_adjoint_precision(_lib.BVE_PRECISION_FP64)


def test_checkpoint_interval_does_not_change_the_gradient(ctx):
    """Recomputation is deterministic: every checkpoint interval gives the
    same bits."""
    g, z = load(ctx, 3)
    F = GaussianWeightFunctional()
    out = [gradient(ctx, F, g.active_area, 0.01, 6, every) for every in (1, 2, 4, 6)]
    for J, gz, gx in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(gz, out[0][1]) and np.array_equal(gx, out[0][2])


def test_repeated_reverse_sweep_gives_the_same_gradient(ctx):
    """The forward sweep keeps the last checkpoint segment's stage inputs,
    which the first reverse sweep uses instead of recomputing them and then
    overwrites with earlier segments; a second reverse sweep after the same
    forward sweep recomputes them and must give the same bits."""
    g, z = load(ctx, 3)
    F = GaussianWeightFunctional()
    xT, zT = ctx.adjoint_forward(0.01, 6, 2)
    lam_x, lam_zeta = F.cotangent(xT, zT, g.active_area)
    first = ctx.adjoint_backward(lam_x, lam_zeta)
    second = ctx.adjoint_backward(lam_x, lam_zeta)
    assert all(np.array_equal(a, b) for a, b in zip(first, second))
    _, gz, gx = gradient(ctx, F, g.active_area, 0.01, 6, 1)
    assert np.array_equal(gz, first[1]) and np.array_equal(gx, first[0])


def test_adjoint_leaves_the_resident_state(ctx):
    g, z = load(ctx, 3)
    before = ctx.store_panels()
    F = GaussianWeightFunctional()
    gradient(ctx, F, g.active_area, 0.01, 3, 2)
    after = ctx.store_panels()
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_adjoint_errors(ctx):
    ctx.reset()
    with pytest.raises(_lib.BVEError, match="no panels loaded"):
        ctx.adjoint_forward(0.01, 2, 1)
    load(ctx, 2)
    ctx.set_precision(_lib.BVE_PRECISION_MIXED)
    try:
        with pytest.raises(_lib.BVEError, match="float64"):
            ctx.adjoint_forward(0.01, 2, 1)
    finally:
        ctx.set_precision(_lib.BVE_PRECISION_FP64)
