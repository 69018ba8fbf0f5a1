"""bve_velocity_noalloc (the allocation-free single evaluation) against
bve_velocity on the same inputs: device-memory inputs and outputs from
torch, two calls in a row with the second smaller, every supported mode;
bit for bit when both run the same launch plan, to the float64 bar on
their default plans; sizes beyond the scratch's maximum and unsupported
modes rejected with an error."""
import numpy as np
import pytest

from thesiscode_amd import _lib
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def noalloc(ctx, t, x, v, a, self_end=0, smooth=(0, 0)):
    """One allocation-free evaluation on device copies of the inputs."""
    dev = lambda arr: torch.from_numpy(np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)).cuda()
    T, X, V, A = dev(t), dev(x), dev(v), dev(a)
    U = torch.full((3 * len(t),), float("nan"), dtype=torch.float64, device="cuda")
    D = torch.full((max(self_end, 1),), float("nan"), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()   # inputs written on torch's stream
    ctx.velocity_noalloc(len(t), T.data_ptr(), len(v), X.data_ptr(), V.data_ptr(), A.data_ptr(),
                         U.data_ptr(), self_end, smooth, D.data_ptr() if self_end else None)
    ctx.synchronize()
    return U.cpu().numpy().reshape(-1, 3), D.cpu().numpy()[:self_end]


def problem(level):
    g = gridmod.cubed_sphere(level)
    z = ic.rh4_wave(g.active_x)
    t = np.concatenate([g.active_x, g.passive_x, g.particles_x])
    return g, z, t, g.n_active, g.n_active + g.n_passive


@pytest.mark.parametrize("precision", [_lib.BVE_PRECISION_FP64, _lib.BVE_PRECISION_FP64_FAST,
                                       _lib.BVE_PRECISION_REFERENCE])
def test_two_calls_match_bve_velocity(precision):
    """A level-5 problem, then a smaller level-3 one on the same scratch
    (stale rows from the first call must not leak into the second), with
    an explicit plan shared by both entry points: bit for bit."""
    with _lib.Context(0) as c:
        c.set_precision(precision)
        c.set_omega(2 * np.pi)
        if precision != _lib.BVE_PRECISION_REFERENCE:
            c.set_plan(0, 512, 384, 0)   # 384 segments: within the scratch's cap
        for level in (5, 3):
            g, z, t, na, nb = problem(level)
            print("level", level, "precision", precision, flush=True)
            ref, ref_dv = c.velocity(t, g.active_x, z, g.active_area, self_end=na,
                                     smooth_range=(na, nb), want_dvort=True)
            got, got_dv = noalloc(c, t, g.active_x, z, g.active_area, na, (na, nb))
            assert np.array_equal(got, ref), level
            assert np.array_equal(got_dv, ref_dv), level


def test_default_plans_agree_at_level7():
    """The level-7 problem (229,376 targets, 98,304 sources) on the default
    plans (the no-allocation one caps the segment count): float64 bar."""
    g, z, t, na, nb = problem(7)
    with _lib.Context(0) as c:
        c.set_precision(_lib.BVE_PRECISION_FP64_FAST)
        ref = c.velocity(t, g.active_x, z, g.active_area, self_end=na, smooth_range=(na, nb))
        got, _ = noalloc(c, t, g.active_x, z, g.active_area, na, (na, nb))
        got2, _ = noalloc(c, t[:5000], g.active_x, z, g.active_area, min(na, 5000))
    scale = np.linalg.norm(ref, axis=1).max()
    assert np.abs(got - ref).max() < 1e-12 * scale
    ref2 = ref[:5000]   # the first 5000 targets are active panels: same sums
    assert np.abs(got2 - ref2).max() < 1e-12 * scale


def test_limits_and_modes_are_checked():
    with _lib.Context(0) as c:
        with pytest.raises(_lib.BVEError, match="exceed the scratch"):
            c.velocity_noalloc(262145, 1, 10, 1, 1, 1, 1)
        c.set_precision(_lib.BVE_PRECISION_MIXED)
        with pytest.raises(_lib.BVEError, match="float64 and reference-order modes"):
            c.velocity_noalloc(10, 1, 10, 1, 1, 1, 1)


def test_profiling_does_not_time_the_noalloc_pass():
    """With profiling on, the allocation-free pass records no timing events
    (they would be created on demand); the result is unchanged."""
    g, z, t, na, nb = problem(3)
    with _lib.Context(0) as c:
        ref, _ = noalloc(c, t, g.active_x, z, g.active_area, na, (na, nb))
        c.set_profiling(True)
        got, _ = noalloc(c, t, g.active_x, z, g.active_area, na, (na, nb))
        launches, _ = c.profile()
        c.set_profiling(False)
    assert launches == 0
    assert np.array_equal(got, ref)
