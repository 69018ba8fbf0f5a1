"""Initial vorticity fields on the sphere (host side).

The velocity kernels do not depend on the initial condition; these only
define the inputs of the benchmark and test workloads.  Shapes follow the
reference's ``TracerAndVorticityDistributions.f90``; the parameters are
arguments so that both the reference's hard-coded values and the benchmark
workloads' values can be produced.

Relative vorticity zeta is what the Biot-Savart sum integrates; absolute
vorticity zeta + 2 Omega z is what a Lagrangian particle carries.
"""
import numpy as np

OMEGA_EARTH = 2.0 * np.pi   # rotation rate per day, omega = 2 PI (TracerAndVorticityDistributions.f90:34-35)


def longitude(x):
    """atan4(y, x) of SphereGeom.f90:213: longitude in [0, 2 PI)."""
    lam = np.arctan2(x[..., 1], x[..., 0])
    return np.where(lam < 0.0, lam + 2.0 * np.pi, lam)


def rh4_wave(x, omega_rh=7.848e-6 * 86400.0, K=7.848e-6 * 86400.0, R=4):
    """Rossby-Haurwitz wave: zeta = 2 w sin(phi) - K (R^2+3R+2) sin(phi) cos^R(phi) cos(R lambda).

    With R = 4 the Legendre factor is z (1 - z^2)^2 and R^2+3R+2 = 30, the
    form of HaurwitzStationary4RelVortVector (TracerAndVorticityDistributions.f90:207-217),
    which hard-codes w = PI/7 and K = -1 (amplitude +30)."""
    z = x[..., 2]
    cosphi2 = 1.0 - z * z
    return (2.0 * omega_rh * z
            - K * (R * R + 3 * R + 2) * z * cosphi2 ** (R // 2) * np.cos(R * longitude(x)))


def rh4_reference(x):
    """The reference's stationary RH4 relative vorticity (alpha = PI/7, +30)."""
    return rh4_wave(x, omega_rh=np.pi / 7.0, K=-1.0)


def latlon_to_xyz(lat, lon):
    return np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])


def gaussian(x, centre, width, amp=4.0 * np.pi):
    """amp * exp(-2 (1 - x.c) / width^2): GeneralGaussian (:893-904) with beta = 1/width."""
    c = np.asarray(centre, dtype=np.float64)
    return amp * np.exp(-2.0 * (1.0 - x @ c) / (width * width))


def zero_mean(zeta_active, area, *others):
    """Subtract the area-weighted mean over the active panels (the gaussConst
    of InitGaussianVortexPanels, :243-276) from every field given."""
    mean = np.sum(zeta_active * area) / (4.0 * np.pi)
    return (zeta_active - mean,) + tuple(o - mean for o in others)


def gauss_vortex(grid, lat0=np.pi / 4, lon0=0.0, width=0.05, amp=4.0 * np.pi):
    """Single Gaussian vortex minus its mean: returns (active, particles, passive) zeta."""
    c = latlon_to_xyz(lat0, lon0)
    za = gaussian(grid.active_x, c, width, amp)
    zp = gaussian(grid.particles_x, c, width, amp)
    zq = gaussian(grid.passive_x, c, width, amp) if grid.n_passive else np.zeros(0)
    return zero_mean(za, grid.active_area, zp, zq)


def two_vortices(grid, lat=np.deg2rad(20.0), lon=np.deg2rad(10.0), width=0.1, amp=4.0 * np.pi):
    """Two co-rotating Gaussians at (lat, +-lon) minus the mean (InitTwoVorts shape, :322-349)."""
    c1, c2 = latlon_to_xyz(lat, lon), latlon_to_xyz(lat, -lon)

    def f(x):
        return gaussian(x, c1, width, amp) + gaussian(x, c2, width, amp)
    zq = f(grid.passive_x) if grid.n_passive else np.zeros(0)
    return zero_mean(f(grid.active_x), grid.active_area, f(grid.particles_x), zq)


def random_field(x, seed, sigma, corr_len=0.3, n_bumps=64):
    """Seeded smooth random perturbation: a sum of n_bumps Gaussian bumps of
    width corr_len at uniformly random centres with N(0, 1) amplitudes,
    scaled so that its RMS over the given points is sigma."""
    rng = np.random.default_rng(seed)
    c = rng.normal(size=(n_bumps, 3))
    c /= np.linalg.norm(c, axis=1, keepdims=True)
    amps = rng.normal(size=n_bumps)
    f = np.exp(-2.0 * (1.0 - x @ c.T) / (corr_len * corr_len)) @ amps
    rms = np.sqrt(np.mean(f * f))
    return f * (sigma / rms) if rms > 0 else f
