"""Host side of the adjoint workload: the Gaussian-weight functional, its
cotangent and its exact difference (CPU only)."""
import numpy as np

from thesiscode_amd.adjoint import GaussianWeightFunctional


def points(n, seed):
    p = np.random.default_rng(seed).normal(size=(n, 3))
    return p / np.linalg.norm(p, axis=1)[:, None]


def test_difference_equals_difference_of_values_at_large_steps():
    """J(x + dx, zeta + dzeta) - J(x, zeta) from the differences equals the
    plain difference where the latter loses little to cancellation."""
    F = GaussianWeightFunctional()
    x = points(500, 0)
    rng = np.random.default_rng(1)
    zeta, area = rng.normal(size=500), rng.uniform(0.5, 1.5, 500) * 4 * np.pi / 500
    dx, dz = 1e-2 * rng.normal(size=(500, 3)), 1e-2 * rng.normal(size=500)
    plain = F.value(x + dx, zeta + dz, area) - F.value(x, zeta, area)
    assert abs(F.difference(x, zeta, dx, dz, area) - plain) <= 1e-12 * abs(plain)


def test_difference_is_exact_at_tiny_steps():
    """At dx, dzeta ~ 1e-12 the difference divided by the step equals the
    directional derivative <cotangent, (dx, dzeta)> to O(step), which a plain
    difference of two values cannot resolve."""
    F = GaussianWeightFunctional()
    x = points(300, 2)
    rng = np.random.default_rng(3)
    zeta, area = rng.normal(size=300), np.full(300, 4 * np.pi / 300)
    vx, vz = rng.normal(size=(300, 3)), rng.normal(size=300)
    h = 1e-12
    lam_x, lam_z = F.cotangent(x, zeta, area)
    deriv = np.sum(lam_x * vx) + np.sum(lam_z * vz)
    got = F.difference(x, zeta, h * vx, h * vz, area) / h
    assert abs(got - deriv) <= 1e-9 * np.sum(np.abs(lam_x * vx)) + 1e-9 * np.sum(np.abs(lam_z * vz))
