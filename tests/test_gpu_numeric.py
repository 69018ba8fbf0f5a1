"""Numeric failure detection: a source and a target at the same point
(1 - x.y = 0 exactly, k != j) make the reference divide 0 by 0
(PanelVelocityParallel.f90:1319-1330, :1366-1369, which only print size
errors, :1300-1315).  libbve flags the non-finite velocity on the device
and the first synchronising call returns BVE_ERR_NUMERIC naming the first
offending target, in every arithmetic mode, instead of carrying NaN
silently; the context stays usable."""
import numpy as np
import pytest

from thesiscode_amd import _lib

pytestmark = pytest.mark.gpu
MODES = [_lib.BVE_PRECISION_FP64, _lib.BVE_PRECISION_FP64_FAST, _lib.BVE_PRECISION_MIXED,
         _lib.BVE_PRECISION_MIXED_KAHAN, _lib.BVE_PRECISION_REFERENCE,
         _lib.BVE_PRECISION_MIXED_BOUNDED]


def panels(n=600, seed=5):
    """n random active panels whose panel 0 sits at the pole (0, 0, 1), a
    point of exactly unit norm, so 1 - x.x is exactly 0 there."""
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(n, 3))
    p /= np.linalg.norm(p, axis=1)[:, None]
    p[0] = (0.0, 0.0, 1.0)
    zeta = rng.normal(size=n)
    area = np.full(n, 4 * np.pi / n)
    return p, zeta, area


@pytest.mark.parametrize("precision", MODES)
def test_particle_on_a_panel_is_reported(precision):
    """A vertex (singular class) at an active panel's centre: the step
    returns BVE_ERR_NUMERIC naming that particle; a clean reload steps."""
    x, zeta, area = panels()
    rng = np.random.default_rng(1)
    q = rng.normal(size=(40, 3))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q[17] = x[0]
    print("mode", precision, flush=True)
    with _lib.Context(0) as c:
        c.set_precision(precision)
        c.load_panels(x, zeta, area, np.zeros((0, 3)))
        with pytest.raises(_lib.BVEError, match="particle 17") as e:
            c.rk4_step(0.01, q)
        assert e.value.code == _lib.BVE_ERR_NUMERIC
        # the status is cleared: a problem without the coincidence steps
        q[17] = -x[0] * 0.6 + np.array([0.8, 0.0, 0.0])
        c.reset()
        c.load_panels(x, zeta, area, np.zeros((0, 3)))
        qn, _, _, _ = c.rk4_step(0.01, q)
        assert np.isfinite(qn).all()


@pytest.mark.parametrize("precision", MODES)
def test_coincident_active_panels_are_reported(precision):
    """Two active panels at the same point (k != j, self-exclusion only
    removes k = j): the resident advance reports the first of them at the
    next synchronisation."""
    x, zeta, area = panels()
    x[5] = x[0]
    print("mode", precision, flush=True)
    with _lib.Context(0) as c:
        c.set_precision(precision)
        c.load_panels(x, zeta, area, np.zeros((0, 3)))
        c.load_particles(np.zeros((0, 3)))
        c.advance(0.01, 1)
        with pytest.raises(_lib.BVEError, match="active panel 0") as e:
            c.synchronize()
        assert e.value.code == _lib.BVE_ERR_NUMERIC
        c.synchronize()   # reported once


@pytest.mark.parametrize("precision", MODES)
def test_single_evaluation_reports_coincident_target(precision):
    """bve_velocity at a target on a source (singular class)."""
    x, zeta, area = panels()
    t = np.array([[1.0, 0.0, 0.0], x[0], [0.0, 1.0, 0.0]])
    with _lib.Context(0) as c:
        c.set_precision(precision)
        with pytest.raises(_lib.BVEError, match="target 1") as e:
            c.velocity(t, x, zeta, area)
        assert e.value.code == _lib.BVE_ERR_NUMERIC
        u = c.velocity(t[[0, 2]], x, zeta, area)
        assert np.isfinite(u).all()


def test_ensemble_reports_member_and_target():
    """Ensemble members share one status word; the report names the member."""
    x, zeta, area = panels(n=300)
    rng = np.random.default_rng(2)
    q = rng.normal(size=(30, 3))
    q /= np.linalg.norm(q, axis=1)[:, None]
    bad = q.copy()
    bad[3] = x[0]
    with _lib.Ensemble(2) as ens:
        ens.set_params(2 * np.pi, 0.05, _lib.BVE_PRECISION_FP64, 0.02)
        ens.load_member(0, x, zeta, area, np.zeros((0, 3)), q, np.zeros(len(q)))
        ens.load_member(1, x, zeta, area, np.zeros((0, 3)), bad, np.zeros(len(q)))
        ens.advance(0.01, 1, use_graph=True)
        with pytest.raises(_lib.BVEError, match="member 1") as e:
            ens.synchronize()
        assert e.value.code == _lib.BVE_ERR_NUMERIC
