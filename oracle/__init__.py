"""ctypes bindings of the CPU oracle (oracle/bve_oracle.c).

TEST INFRASTRUCTURE: only tests/, the smoke check and the benchmark's CPU
baseline leg import this package, as the checker.  The product library
(thesiscode_amd) never imports it.

Arrays are float64 numpy arrays; positions are (N, 3) C-contiguous, i.e.
the reference's column-major x(3, N).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_d = ctypes.c_double
_i = ctypes.c_int
_p = ctypes.c_void_p


def build():
    """Compile liboracle.so (gcc, OpenMP)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.bvo_active_rhs.argtypes = [_i, _p, _p, _p, _i, _i, _d, _p, _p]
        L.bvo_passive_rhs.argtypes = [_i, _p, _i, _p, _p, _p, _i, _i, _p]
        L.bvo_smooth_velocity.argtypes = [_i, _p, _i, _p, _p, _p, _i, _i, _d, _p]
        L.bvo_voronoi_velocity.argtypes = [_i, _p, _p, _p, _p, _p, _i, _i, _d, _p, _p, _p, _p]
        L.bvo_rk4_step.argtypes = [_i, _p, _p, _p, _i, _p, _i, _p, _p, _p, _d, _d, _d, _p, _p]
        L.bvo_voronoi_rk4_step.argtypes = [_i, _p, _p, _p, _p, _p, _p, _d, _d, _p, _p]
        L.bvo_renormalize.argtypes = [_i, _p]
        L.bvo_wind.argtypes = [_i, _p, _d, _p]
        L.bvo_advection_rk4_step.argtypes = [_i, _i, _p, _d, _d]
        L.bvo_juckes_forcing.argtypes = [_p, _d]
        L.bvo_juckes_forcing.restype = _d
        L.bvo_stratosphere_rk4_step.argtypes = [_i, _p, _p, _p, _p, _i, _p, _i, _p, _p, _p, _d,
                                                _d, _d, _d, _p, _p]
        L.bvo_velocity_q.argtypes = [_i, _p, _i, _p, _p, _i, _d, _p]
        L.bvo_jvp_active.argtypes = [_i, _p, _p, _p, _p, _p, _d, _p, _p]
        L.bvo_tlm_rk4.argtypes = [_i, _p, _p, _p, _p, _p, _d, _d, _i]
        L.bvo_num_threads.restype = _i
        L.bvo_set_num_threads.argtypes = [_i]
        _lib = L
    return _lib


def _c(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _ptr(a):
    return a.ctypes.data_as(_p) if a is not None else None


def active_rhs(x, vort, area, omega, i0=0, i1=None):
    """BVEActiveRHS: returns (dX (N,3), dVort (N,))."""
    x, vort, area = _c(x), _c(vort), _c(area)
    n = len(vort)
    i1 = n if i1 is None else i1
    dx = np.zeros((n, 3))
    dv = np.zeros(n)
    lib().bvo_active_rhs(n, _ptr(x), _ptr(vort), _ptr(area), i0, i1, omega, _ptr(dx), _ptr(dv))
    return dx, dv


def passive_rhs(X, x, vort, area, i0=0, i1=None):
    """BVEPassiveRHS: singular sum at targets X."""
    X, x, vort, area = _c(X), _c(x), _c(vort), _c(area)
    m = len(X)
    i1 = m if i1 is None else i1
    dx = np.zeros((m, 3))
    lib().bvo_passive_rhs(m, _ptr(X), len(vort), _ptr(x), _ptr(vort), _ptr(area), i0, i1, _ptr(dx))
    return dx


def smooth_velocity(X, x, vort, area, smooth, i0=0, i1=None):
    """BVESmoothVelocity: regularised sum at targets X."""
    X, x, vort, area = _c(X), _c(x), _c(vort), _c(area)
    m = len(X)
    i1 = m if i1 is None else i1
    dx = np.zeros((m, 3))
    lib().bvo_smooth_velocity(m, _ptr(X), len(vort), _ptr(x), _ptr(vort), _ptr(area), i0, i1,
                              smooth, _ptr(dx))
    return dx


def voronoi_velocity(x, y, z, vort, area, omega, i0=0, i1=None):
    """Voronoi BVEVelocity (SoA): returns (dx, dy, dz, dVort)."""
    x, y, z, vort, area = (_c(a) for a in (x, y, z, vort, area))
    n = len(x)
    i1 = n if i1 is None else i1
    out = [np.zeros(n) for _ in range(4)]
    lib().bvo_voronoi_velocity(n, _ptr(x), _ptr(y), _ptr(z), _ptr(vort), _ptr(area), i0, i1,
                               omega, *(_ptr(o) for o in out))
    return tuple(out)


def rk4_step(active_x, active_vort, area, passive_x, particles_x, particles_absvort,
             dt, omega, smooth):
    """One BVERK4 step.  Returns a dict with the new state, the stage-1
    kinetic energy per active panel and totalKE."""
    ax, av, ar = _c(active_x).copy(), _c(active_vort).copy(), _c(area)
    px, qx = _c(passive_x).copy(), _c(particles_x).copy()
    qabs = _c(particles_absvort)
    na, npp, np_ = len(av), len(px), len(qx)
    qrel = np.zeros(np_)
    ke = np.zeros(na)
    tot = ctypes.c_double(0.0)
    lib().bvo_rk4_step(na, _ptr(ax), _ptr(av), _ptr(ar), npp, _ptr(px), np_, _ptr(qx), _ptr(qrel),
                       _ptr(qabs), dt, omega, smooth, _ptr(ke), ctypes.byref(tot))
    return dict(active_x=ax, active_relvort=av, passive_x=px, particles_x=qx,
                particles_relvort=qrel, ke=ke, total_ke=tot.value)


def voronoi_rk4_step(x, y, z, relvort, sum_area, cur_area, dt, omega):
    """One Voronoi BVERK4 step (SoA) with NormalizeGrid.  sum_area weighs the
    sources, cur_area enters only totalKE.  Returns (x, y, z, relvort, ke,
    total_ke)."""
    x, y, z, rv = (_c(a).copy() for a in (x, y, z, relvort))
    sa, ca = _c(sum_area), _c(cur_area)
    n = len(x)
    ke = np.zeros(n)
    tot = ctypes.c_double(0.0)
    lib().bvo_voronoi_rk4_step(n, _ptr(x), _ptr(y), _ptr(z), _ptr(rv), _ptr(sa), _ptr(ca), dt,
                               omega, _ptr(ke), ctypes.byref(tot))
    return x, y, z, rv, ke, tot.value


def wind(x, t, divergent=False):
    """NonDivergentWind / DivergentWind at points x (N, 3) and time t."""
    x = _c(x).reshape(-1, 3)
    out = np.zeros_like(x)
    for i in range(len(x)):
        lib().bvo_wind(int(divergent), _ptr(np.ascontiguousarray(x[i])), t, _ptr(out[i]))
    return out


def advection_rk4_step(x, t, dt, divergent=False):
    """One AdvectionRK4 / DivergentAdvectionRK4 step of the points x (N, 3)."""
    x = _c(x).copy()
    lib().bvo_advection_rk4_step(int(divergent), len(x), _ptr(x), t, dt)
    return x


def juckes_forcing(x, t):
    """Juckes_Forcing at points x (N, 3) and time t."""
    x = _c(x).reshape(-1, 3)
    return np.array([lib().bvo_juckes_forcing(_ptr(np.ascontiguousarray(p)), t) for p in x])


def stratosphere_rk4_step(active_x, active_absvort, area, passive_x, particles_x,
                          particles_absvort, t, dt, omega, smooth):
    """One StratosphereRK4 step from time t.  Returns a dict like rk4_step's;
    active_relvort and particles_relvort are absVort - 2 Omega z - F(x, t+dt)."""
    ax, aabs, ar = _c(active_x).copy(), _c(active_absvort), _c(area)
    px, qx = _c(passive_x).copy(), _c(particles_x).copy()
    qabs = _c(particles_absvort)
    na, npp, np_ = len(aabs), len(px), len(qx)
    av, qrel, ke = np.zeros(na), np.zeros(np_), np.zeros(na)
    tot = ctypes.c_double(0.0)
    lib().bvo_stratosphere_rk4_step(na, _ptr(ax), _ptr(av), _ptr(aabs), _ptr(ar), npp, _ptr(px),
                                    np_, _ptr(qx), _ptr(qrel), _ptr(qabs), t, dt, omega, smooth,
                                    _ptr(ke), ctypes.byref(tot))
    return dict(active_x=ax, active_relvort=av, passive_x=px, particles_x=qx,
                particles_relvort=qrel, ke=ke, total_ke=tot.value)


def renormalize(x):
    x = _c(x).copy()
    lib().bvo_renormalize(len(x), _ptr(x))
    return x


def velocity_q(X, x, w, self_off=-1, eps2=0.0):
    """__float128 brute force: u at targets X from sources x with weights w."""
    X, x, w = _c(X), _c(x), _c(w)
    u = np.zeros((len(X), 3))
    lib().bvo_velocity_q(len(X), _ptr(X), len(w), _ptr(x), _ptr(w), self_off, eps2, _ptr(u))
    return u


def jvp_active(x, zeta, area, dx, dzeta, omega):
    """Jacobian-vector product of the active-panel right-hand side."""
    x, zeta, area, dx, dzeta = (_c(a) for a in (x, zeta, area, dx, dzeta))
    n = len(zeta)
    du = np.zeros((n, 3))
    dz = np.zeros(n)
    lib().bvo_jvp_active(n, _ptr(x), _ptr(zeta), _ptr(area), _ptr(dx), _ptr(dzeta), omega,
                         _ptr(du), _ptr(dz))
    return du, dz


def tlm_rk4(x, zeta, area, dx, dzeta, dt, omega, n_steps):
    """Tangent-linear model of n_steps RK4 steps of the active panels.
    Returns (x_T, zeta_T, dx_T, dzeta_T)."""
    x, zeta, dx, dzeta = (_c(a).copy() for a in (x, zeta, dx, dzeta))
    area = _c(area)
    lib().bvo_tlm_rk4(len(zeta), _ptr(x), _ptr(zeta), _ptr(area), _ptr(dx), _ptr(dzeta), dt, omega,
                      n_steps)
    return x, zeta, dx, dzeta


def num_threads():
    return lib().bvo_num_threads()


def set_num_threads(n):
    lib().bvo_set_num_threads(n)
