"""ctypes binding of libbve.so (the HIP library) and a thin numpy wrapper.

The library must have been built (thesiscode_amd.build.build()); loading
fails loudly if it is missing -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbve.so")

BVE_OK = 0
BVE_ERR_NUMERIC = 6
BVE_PRECISION_FP64 = 0
BVE_PRECISION_MIXED = 1
BVE_PRECISION_FP64_FAST = 2
BVE_PRECISION_MIXED_KAHAN = 3
BVE_PRECISION_REFERENCE = 4
BVE_PRECISION_MIXED_BOUNDED = 5

# every symbol include/bve.h declares, with its ctypes signature
_d, _i, _p, _ll = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "bve_create": (_i, [ctypes.POINTER(_p), _i]),
    "bve_destroy": (_i, [_p]),
    "bve_last_error": (ctypes.c_char_p, [_p]),
    "bve_set_omega": (_i, [_p, _d]),
    "bve_set_smooth": (_i, [_p, _d]),
    "bve_get_smooth": (_i, [_p, _dp]),
    "bve_set_precision": (_i, [_p, _i, _d]),
    "bve_set_renormalize": (_i, [_p, _i]),
    "bve_set_error_target": (_i, [_p, _d]),
    "bve_set_plan": (_i, [_p, _i, _i, _i, _i]),
    "bve_set_chunk": (_i, [_p, _i]),
    "bve_load_panels": (_i, [_p, _i, _p, _p, _p, _i, _p]),
    "bve_rk4_step": (_i, [_p, _d, _i, _p, _p, _p, _p, _p]),
    "bve_load_panel_absvort": (_i, [_p, _p]),
    "bve_advection_rk4_step": (_i, [_p, _d, _d, _i, _i, _p]),
    "bve_stratosphere_rk4_step": (_i, [_p, _d, _d, _i, _p, _p, _p, _p, _p]),
    "bve_store_panels": (_i, [_p, _p, _p, _p]),
    "bve_reset": (_i, [_p]),
    "bve_load_particles": (_i, [_p, _i, _p, _p]),
    "bve_advance": (_i, [_p, _d, _i]),
    "bve_store_particles": (_i, [_p, _p, _p]),
    "bve_get_kinetic_energy": (_i, [_p, _p, _p]),
    "bve_synchronize": (_i, [_p]),
    "bve_velocity": (_i, [_p, _i, _p, _i, _i, _i, _i, _p, _p, _p, _p, _p]),
    "bve_resident_velocity": (_i, [_p, _p, _p, _p]),
    "bve_velocity_noalloc": (_i, [_p, _i, _p, _i, _i, _i, _i, _p, _p, _p, _p, _p]),
    "bve_voronoi_rk4_step": (_i, [_p, _d, _i, _p, _p, _p, _p, _p, _p, _p]),
    "bve_difference_forward": (_i, [_p, _d, _i, _p, _p, _p, _p, _p]),
    "bve_adjoint_forward": (_i, [_p, _d, _i, _i, _p, _p]),
    "bve_adjoint_backward": (_i, [_p, _p, _p, _p, _p]),
    "bve_set_adjoint_precision": (_i, [_p, _i, _d]),
    "bve_ensemble_create": (_i, [ctypes.POINTER(_p), _i, _i]),
    "bve_ensemble_destroy": (_i, [_p]),
    "bve_ensemble_last_error": (ctypes.c_char_p, [_p]),
    "bve_ensemble_set_params": (_i, [_p, _d, _d, _i, _d]),
    "bve_ensemble_load_member": (_i, [_p, _i, _i, _p, _p, _p, _i, _p, _i, _p, _p]),
    "bve_ensemble_set_member_plan": (_i, [_p, _i, _i, _i]),
    "bve_ensemble_get_member_plan": (_i, [_p, _i, _p, _p, _p]),
    "bve_ensemble_set_items_per_workgroup": (_i, [_p, _i]),
    "bve_ensemble_set_grid": (_i, [_p, _i]),
    "bve_ensemble_advance": (_i, [_p, _d, _i, _i]),
    "bve_ensemble_synchronize": (_i, [_p]),
    "bve_ensemble_store_member": (_i, [_p, _i, _p, _p, _p, _p, _p, _p]),
    "bve_ensemble_set_profiling": (_i, [_p, _i]),
    "bve_ensemble_get_profile": (_i, [_p, ctypes.POINTER(_ll), _dp]),
    "bve_ensemble_pairs_per_step": (_i, [_p, _dp]),
    "bve_set_profiling": (_i, [_p, _i]),
    "bve_get_profile": (_i, [_p, ctypes.POINTER(_ll), _dp]),
    "bve_get_plan": (_i, [_p, _p, _p, _p, _p, _p]),
    "bve_get_launch": (_i, [_p, _p, _p, _p, _p]),
    "bve_pairs_per_step": (_i, [_p, _dp]),
    "bve_vtk_write_frame": (_i, [ctypes.c_char_p, _i, _i, _p, _p, _p, _i, _p, _i, _i, _i, _i,
                                 _p, _p, _p, _p, _p]),
}

_lib = None


class BVEError(RuntimeError):
    """A libbve call failed; ``code`` is its status (BVE_ERR_*)."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


def load():
    """Load libbve.so and declare its signatures (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BVEError(f"{LIB_PATH} is not built: run thesiscode_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_p)


class Context:
    """One libbve context (device state of one solver instance)."""

    def __init__(self, device=0):
        L = load()
        h = _p()
        rc = L.bve_create(ctypes.byref(h), device)
        if rc != BVE_OK:
            raise BVEError(f"bve_create failed ({rc}): {L.bve_last_error(None).decode()}")
        self._h = h
        self._L = L

    def close(self):
        if getattr(self, "_h", None):
            self._L.bve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != BVE_OK:
            raise BVEError(f"{what} failed ({rc}): {self._L.bve_last_error(self._h).decode()}",
                           rc)

    # parameters
    def set_omega(self, omega):
        self._check(self._L.bve_set_omega(self._h, omega), "bve_set_omega")

    def set_smooth(self, smooth):
        self._check(self._L.bve_set_smooth(self._h, smooth), "bve_set_smooth")

    def get_smooth(self):
        v = ctypes.c_double()
        self._check(self._L.bve_get_smooth(self._h, ctypes.byref(v)), "bve_get_smooth")
        return v.value

    def set_precision(self, precision, nearfield_radius=0.02):
        self._check(self._L.bve_set_precision(self._h, precision, nearfield_radius),
                    "bve_set_precision")

    def set_error_target(self, rel_err):
        self._check(self._L.bve_set_error_target(self._h, rel_err), "bve_set_error_target")

    def set_renormalize(self, on):
        self._check(self._L.bve_set_renormalize(self._h, int(bool(on))), "bve_set_renormalize")

    def set_plan(self, targets_per_lane=0, tile=0, target_workgroups=0, recip=0):
        self._check(self._L.bve_set_plan(self._h, targets_per_lane, tile, target_workgroups, recip),
                    "bve_set_plan")

    def set_chunk(self, chunk=0):
        self._check(self._L.bve_set_chunk(self._h, chunk), "bve_set_chunk")

    # panel path
    def load_panels(self, active_x, active_relvort, active_area, passive_x):
        na = len(active_relvort)
        npp = len(passive_x)
        ax = _f64(active_x, 3 * na)
        av = _f64(active_relvort, na)
        ar = _f64(active_area, na)
        px = _f64(passive_x, 3 * npp) if npp else None
        self._check(self._L.bve_load_panels(self._h, na, _ptr(ax), _ptr(av), _ptr(ar), npp, _ptr(px)),
                    "bve_load_panels")
        self.n_active, self.n_passive = na, npp
        self.n_particles = 0

    def rk4_step(self, dt, particles_x, absvort=None, want_ke=True):
        """One step; returns (new particle x, particle relvort or None, totalKE, ke)."""
        n = len(particles_x)
        qx = _f64(particles_x, 3 * n).copy()
        qa = _f64(absvort, n) if absvort is not None else None
        rel = np.zeros(n) if absvort is not None else None
        tot = np.zeros(1)
        ke = np.zeros(self.n_active) if want_ke else None
        self._check(self._L.bve_rk4_step(self._h, dt, n, _ptr(qx), _ptr(qa), _ptr(rel), _ptr(tot),
                                         _ptr(ke)), "bve_rk4_step")
        self.n_particles = n
        return qx.reshape(n, 3), rel, float(tot[0]), ke

    def rk4_step_into(self, dt, particles_x, absvort, relvort, ke):
        """rk4_step on caller-owned arrays (the Fortran drop-in's pattern):
        particles_x (n, 3) float64 C-contiguous is updated in place, relvort
        (n) and ke (n_active) are written; returns totalKE."""
        n = len(particles_x)
        for a, size in ((particles_x, 3 * n), (absvort, n), (relvort, n), (ke, self.n_active)):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.size != size:
                raise ValueError("rk4_step_into needs C-contiguous float64 arrays of the right size")
        tot = np.zeros(1)
        self._check(self._L.bve_rk4_step(self._h, dt, n, _ptr(particles_x), _ptr(absvort),
                                         _ptr(relvort), _ptr(tot), _ptr(ke)), "bve_rk4_step")
        self.n_particles = n
        return float(tot[0])

    def store_panels_into(self, active_x, active_relvort, passive_x):
        """store_panels into caller-owned C-contiguous float64 arrays."""
        for a, size in ((active_x, 3 * self.n_active), (active_relvort, self.n_active),
                        (passive_x, 3 * self.n_passive)):
            if a.dtype != np.float64 or not a.flags.c_contiguous or a.size != size:
                raise ValueError("store_panels_into needs C-contiguous float64 arrays")
        self._check(self._L.bve_store_panels(self._h, _ptr(active_x), _ptr(active_relvort),
                                             _ptr(passive_x)), "bve_store_panels")

    def advection_rk4_step(self, t, dt, particles_x, divergent=False):
        """One AdvectionRK4 / DivergentAdvectionRK4 step; returns the particles."""
        n = len(particles_x)
        qx = _f64(particles_x, 3 * n).copy()
        self._check(self._L.bve_advection_rk4_step(self._h, t, dt, int(divergent), n, _ptr(qx)),
                    "bve_advection_rk4_step")
        self.n_particles = n
        return qx.reshape(n, 3)

    def load_panel_absvort(self, absvort_active):
        a = _f64(absvort_active, self.n_active)
        self._check(self._L.bve_load_panel_absvort(self._h, _ptr(a)), "bve_load_panel_absvort")

    def stratosphere_rk4_step(self, t, dt, particles_x, absvort=None, want_ke=True):
        """One forced (StratosphereRK4) step from time t; returns like rk4_step."""
        n = len(particles_x)
        qx = _f64(particles_x, 3 * n).copy()
        qa = _f64(absvort, n) if absvort is not None else None
        rel = np.zeros(n) if absvort is not None else None
        tot = np.zeros(1)
        ke = np.zeros(self.n_active) if want_ke else None
        self._check(self._L.bve_stratosphere_rk4_step(self._h, t, dt, n, _ptr(qx), _ptr(qa),
                                                      _ptr(rel), _ptr(tot), _ptr(ke)),
                    "bve_stratosphere_rk4_step")
        self.n_particles = n
        return qx.reshape(n, 3), rel, float(tot[0]), ke

    def store_panels(self):
        ax = np.zeros((self.n_active, 3))
        av = np.zeros(self.n_active)
        px = np.zeros((self.n_passive, 3))
        self._check(self._L.bve_store_panels(self._h, _ptr(ax), _ptr(av), _ptr(px)),
                    "bve_store_panels")
        return ax, av, px

    def reset(self):
        self._check(self._L.bve_reset(self._h), "bve_reset")

    # resident stepping
    def load_particles(self, particles_x, absvort=None):
        n = len(particles_x)
        qx = _f64(particles_x, 3 * n)
        qa = _f64(absvort, n) if absvort is not None else None
        self._check(self._L.bve_load_particles(self._h, n, _ptr(qx), _ptr(qa)), "bve_load_particles")
        self.n_particles = n

    def advance(self, dt, n_steps):
        self._check(self._L.bve_advance(self._h, dt, n_steps), "bve_advance")

    def store_particles(self, want_relvort=False):
        qx = np.zeros((self.n_particles, 3))
        rel = np.zeros(self.n_particles) if want_relvort else None
        self._check(self._L.bve_store_particles(self._h, _ptr(qx), _ptr(rel)), "bve_store_particles")
        return qx, rel

    def kinetic_energy(self):
        tot = np.zeros(1)
        ke = np.zeros(self.n_active)
        self._check(self._L.bve_get_kinetic_energy(self._h, _ptr(tot), _ptr(ke)),
                    "bve_get_kinetic_energy")
        return float(tot[0]), ke

    def synchronize(self):
        self._check(self._L.bve_synchronize(self._h), "bve_synchronize")

    def resident_velocity(self):
        """Velocities (u_active, u_passive, u_particles) of the resident state,
        evaluated with the context's own launch plan and arithmetic."""
        ua = np.zeros((self.n_active, 3))
        up = np.zeros((self.n_passive, 3))
        uq = np.zeros((self.n_particles, 3))
        self._check(self._L.bve_resident_velocity(self._h, _ptr(ua), _ptr(up), _ptr(uq)),
                    "bve_resident_velocity")
        return ua, up, uq

    def velocity_noalloc(self, n_targets, x_targets, n_sources, x_sources, vort, area, u,
                         self_end=0, smooth_range=(0, 0), dvort=None):
        """bve_velocity_noalloc on device addresses (ints): enqueued on the
        context's stream, no allocation, no wait (synchronize() waits)."""
        self._check(self._L.bve_velocity_noalloc(self._h, n_targets, x_targets, self_end,
                                                 smooth_range[0], smooth_range[1], n_sources,
                                                 x_sources, vort, area, u, dvort),
                    "bve_velocity_noalloc")

    # single evaluation
    def velocity(self, targets_x, sources_x, vort, area, self_end=0, smooth_range=(0, 0),
                 want_dvort=False):
        nt, ns = len(targets_x), len(vort)
        tx = _f64(targets_x, 3 * nt)
        sx = _f64(sources_x, 3 * ns)
        v = _f64(vort, ns)
        a = _f64(area, ns)
        u = np.zeros((nt, 3))
        dv = np.zeros(self_end) if want_dvort else None
        self._check(self._L.bve_velocity(self._h, nt, _ptr(tx), self_end, smooth_range[0],
                                         smooth_range[1], ns, _ptr(sx), _ptr(v), _ptr(a), _ptr(u),
                                         _ptr(dv)), "bve_velocity")
        return (u, dv) if want_dvort else u

    # Voronoi path
    def voronoi_rk4_step(self, dt, x, y, z, relvort, area):
        n = len(x)
        xs = [_f64(a, n).copy() for a in (x, y, z, relvort)]
        ar = _f64(area, n)
        tot = np.zeros(1)
        ke = np.zeros(n)
        self._check(self._L.bve_voronoi_rk4_step(self._h, dt, n, *(_ptr(a) for a in xs), _ptr(ar),
                                                 _ptr(tot), _ptr(ke)), "bve_voronoi_rk4_step")
        return xs[0], xs[1], xs[2], xs[3], float(tot[0]), ke

    # exact differences
    def difference_forward(self, dt, n_steps, dzeta0):
        """Base trajectory and its exact difference to the trajectory with
        vorticity zeta0 + dzeta0 (active panels); returns (x_T, zeta_T, dx_T,
        dzeta_T)."""
        n = self.n_active
        dz0 = _f64(dzeta0, n)
        xT, zT, dxT, dzT = np.zeros((n, 3)), np.zeros(n), np.zeros((n, 3)), np.zeros(n)
        self._check(self._L.bve_difference_forward(self._h, dt, n_steps, _ptr(dz0), _ptr(xT),
                                                   _ptr(zT), _ptr(dxT), _ptr(dzT)),
                    "bve_difference_forward")
        return xT, zT, dxT, dzT

    # adjoint
    def set_adjoint_precision(self, precision, ratio=0.7):
        """Arithmetic of the adjoint's pair pass: BVE_PRECISION_FP64 or
        BVE_PRECISION_MIXED (float32 for well-separated group/tile pairs)."""
        self._check(self._L.bve_set_adjoint_precision(self._h, precision, ratio),
                    "bve_set_adjoint_precision")

    def adjoint_forward(self, dt, n_steps, checkpoint_every):
        """Active-panel forward sweep with checkpoints; returns (x_T, zeta_T)."""
        xT = np.zeros((self.n_active, 3))
        zT = np.zeros(self.n_active)
        self._check(self._L.bve_adjoint_forward(self._h, dt, n_steps, checkpoint_every, _ptr(xT),
                                                _ptr(zT)), "bve_adjoint_forward")
        return xT, zT

    def adjoint_backward(self, lambda_x, lambda_zeta):
        """Reverse sweep; returns (dJ/dx_0, dJ/dzeta_0)."""
        n = self.n_active
        lx = _f64(lambda_x, 3 * n)
        lz = _f64(lambda_zeta, n)
        gx = np.zeros((n, 3))
        gz = np.zeros(n)
        self._check(self._L.bve_adjoint_backward(self._h, _ptr(lx), _ptr(lz), _ptr(gx), _ptr(gz)),
                    "bve_adjoint_backward")
        return gx, gz

    # measurement
    def set_profiling(self, on):
        self._check(self._L.bve_set_profiling(self._h, int(bool(on))), "bve_set_profiling")

    def profile(self):
        n = ctypes.c_longlong()
        ms = ctypes.c_double()
        self._check(self._L.bve_get_profile(self._h, ctypes.byref(n), ctypes.byref(ms)),
                    "bve_get_profile")
        return n.value, ms.value

    def pairs_per_step(self):
        v = ctypes.c_double()
        self._check(self._L.bve_pairs_per_step(self._h, ctypes.byref(v)), "bve_pairs_per_step")
        return v.value


    def plan(self):
        """(targets_per_lane, tile, chunk, n_chunks, recip) of the pair kernel."""
        v = [ctypes.c_int() for _ in range(5)]
        self._check(self._L.bve_get_plan(self._h, *(ctypes.byref(x) for x in v)), "bve_get_plan")
        return tuple(x.value for x in v)

    def launch(self):
        """{segments, grid, n_tiles, partial_bytes} of the pair kernel."""
        v = [ctypes.c_int() for _ in range(3)]
        pb = ctypes.c_longlong()
        self._check(self._L.bve_get_launch(self._h, *(ctypes.byref(x) for x in v), ctypes.byref(pb)),
                    "bve_get_launch")
        return dict(segments=v[0].value, grid=v[1].value, n_tiles=v[2].value,
                    partial_bytes=pb.value)


class Ensemble:
    """Many independent panel problems advanced by shared launches."""

    def __init__(self, n_members, device=0):
        L = load()
        h = _p()
        rc = L.bve_ensemble_create(ctypes.byref(h), device, n_members)
        if rc != BVE_OK:
            raise BVEError(f"bve_ensemble_create failed ({rc}): "
                           f"{L.bve_ensemble_last_error(None).decode()}")
        self._h, self._L = h, L
        self.n_members = n_members
        self._sizes = {}

    def close(self):
        if getattr(self, "_h", None):
            self._L.bve_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc, what):
        if rc != BVE_OK:
            raise BVEError(f"{what} failed ({rc}): {self._L.bve_ensemble_last_error(self._h).decode()}",
                           rc)

    def set_params(self, omega, smooth=0.05, precision=BVE_PRECISION_FP64, nearfield_radius=0.02):
        self._check(self._L.bve_ensemble_set_params(self._h, omega, smooth, precision,
                                                    nearfield_radius), "bve_ensemble_set_params")

    def load_member(self, m, active_x, active_relvort, active_area, passive_x, particles_x,
                    particles_absvort=None):
        na, npp, npt = len(active_relvort), len(passive_x), len(particles_x)
        ax, av, ar = _f64(active_x, 3 * na), _f64(active_relvort, na), _f64(active_area, na)
        px = _f64(passive_x, 3 * npp) if npp else None
        qx = _f64(particles_x, 3 * npt) if npt else None
        qa = _f64(particles_absvort, npt) if particles_absvort is not None else None
        self._check(self._L.bve_ensemble_load_member(self._h, m, na, _ptr(ax), _ptr(av), _ptr(ar),
                                                     npp, _ptr(px), npt, _ptr(qx), _ptr(qa)),
                    "bve_ensemble_load_member")
        self._sizes[m] = (na, npp, npt)

    def set_member_plan(self, m, targets_per_lane=0, chunk=0):
        self._check(self._L.bve_ensemble_set_member_plan(self._h, m, targets_per_lane, chunk),
                    "bve_ensemble_set_member_plan")

    def member_plan(self, m):
        """(targets per lane, sources per work item, work items) of member m
        as the next advance runs it."""
        t, c, w = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self._L.bve_ensemble_get_member_plan(self._h, m, ctypes.byref(t),
                                                         ctypes.byref(c), ctypes.byref(w)),
                    "bve_ensemble_get_member_plan")
        return t.value, c.value, w.value

    def set_items_per_workgroup(self, items):
        self._check(self._L.bve_ensemble_set_items_per_workgroup(self._h, items),
                    "bve_ensemble_set_items_per_workgroup")

    def set_grid(self, workgroups):
        self._check(self._L.bve_ensemble_set_grid(self._h, workgroups), "bve_ensemble_set_grid")

    def advance(self, dt, n_steps, use_graph=True):
        self._check(self._L.bve_ensemble_advance(self._h, dt, n_steps, int(bool(use_graph))),
                    "bve_ensemble_advance")

    def synchronize(self):
        self._check(self._L.bve_ensemble_synchronize(self._h), "bve_ensemble_synchronize")

    def store_member(self, m):
        na, npp, npt = self._sizes[m]
        ax, av, px = np.zeros((na, 3)), np.zeros(na), np.zeros((npp, 3))
        qx, qr, tot = np.zeros((npt, 3)), np.zeros(npt), np.zeros(1)
        self._check(self._L.bve_ensemble_store_member(self._h, m, _ptr(ax), _ptr(av), _ptr(px),
                                                      _ptr(qx), None, _ptr(tot)),
                    "bve_ensemble_store_member")
        return dict(active_x=ax, active_relvort=av, passive_x=px, particles_x=qx,
                    total_ke=float(tot[0]))

    def set_profiling(self, on):
        self._check(self._L.bve_ensemble_set_profiling(self._h, int(bool(on))),
                    "bve_ensemble_set_profiling")

    def profile(self):
        n = ctypes.c_longlong()
        ms = ctypes.c_double()
        self._check(self._L.bve_ensemble_get_profile(self._h, ctypes.byref(n), ctypes.byref(ms)),
                    "bve_ensemble_get_profile")
        return n.value, ms.value

    def pairs_per_step(self):
        v = ctypes.c_double()
        self._check(self._L.bve_ensemble_pairs_per_step(self._h, ctypes.byref(v)),
                    "bve_ensemble_pairs_per_step")
        return v.value
