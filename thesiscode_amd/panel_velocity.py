"""Host-side mirror of the reference's velocity modules, backed by libbve.

``PanelVelocity`` has the public interface of ``PanelVelocityParallelModule``
(PanelVelocityParallel.f90:14-20): BVERK4, ResetRK4, InitializeMPIRK4,
FinalizeMPIRK4, LoadBalance, SetSmooth, GetSmooth, ResetArea, totalKE and
kineticEnergy, with the same argument meaning and the same caching rules:

* the active and passive panels are gathered once after a reset (InitRK4 +
  GatherPanels, :236-318) and the device copy is authoritative until
  ResetRK4 -- host edits to panel positions or vorticity in between are
  ignored and overwritten by the scatter after each step (:656);
* particle positions are read from the host at every step (:458) and the
  particles' relative vorticity is reset from their absolute vorticity
  after the step, relVort = absVort - 2 Omega z (:651);
* kineticEnergy (|u|^2 at stage 1) and totalKE (:495-501) are set by every
  step and written to panel tracer 2 when there are >= 2 tracers.

procRank / numProcs are accepted and do not change results (one GPU does
all targets); LoadBalance still returns the reference's partition.
``VoronoiVelocity`` mirrors VoronoiVelocityParallel.f90's BVERK4 on SoA
generators.

``Particles`` / ``Panels`` hold the fields of the reference's types that the
path reads or writes (ParticlesEdgesPanels2.f90:45-89), with positions as
(N, 3) arrays (row j = column j of the Fortran x(3, N)).
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import _lib
from .initial_conditions import OMEGA_EARTH


@dataclass
class Particles:
    x: np.ndarray                       # (N, 3)
    relVort: np.ndarray                 # (N,)
    absVort: np.ndarray                 # (N,)
    x0: Optional[np.ndarray] = None
    tracer: Optional[np.ndarray] = None  # (N, nTracer)

    @property
    def N(self):
        return self.x.shape[0]


@dataclass
class Panels:
    x: np.ndarray                       # (N, 3) all panels, tree order
    area: np.ndarray                    # (N,) zero for divided panels
    hasChildren: np.ndarray             # (N,) bool
    relVort: np.ndarray                 # (N,)
    absVort: np.ndarray                 # (N,)
    x0: Optional[np.ndarray] = None
    tracer: Optional[np.ndarray] = None  # (N, nTracer)
    vertices: Optional[np.ndarray] = None  # (N, panelKind) 0-based particle indices (corners)
    panelKind: int = 4                     # QUAD_PANEL; 3 = TRI_PANEL

    @property
    def N(self):
        return self.x.shape[0]

    @property
    def N_Active(self):
        return int(np.count_nonzero(~self.hasChildren))


def panels_from_grid(grid, relvort_active, relvort_passive=None, omega=OMEGA_EARTH, n_tracer=3):
    """Panels of a uniform cubed sphere (thesiscode_amd.grid) in tree order:
    the passive (divided) panels of levels 0..L-1 come first in the
    reference's numbering, then the leaves; GatherPanels only needs the
    active/passive split, which this preserves.  Corner lists are kept for
    the leaves only (-1 for divided panels, which no frame writes)."""
    na, npp = grid.n_active, grid.n_passive
    x = np.concatenate([grid.passive_x, grid.active_x])
    area = np.concatenate([np.zeros(npp), grid.active_area])
    has = np.concatenate([np.ones(npp, bool), np.zeros(na, bool)])
    rv = np.concatenate([np.zeros(npp) if relvort_passive is None else relvort_passive,
                         relvort_active])
    verts = np.concatenate([np.full((npp, 4), -1, np.int64), grid.active_verts])
    return Panels(x=x, area=area, hasChildren=has, relVort=rv, absVort=rv + 2.0 * omega * x[:, 2],
                  x0=x.copy(), tracer=np.zeros((na + npp, n_tracer)), vertices=verts)


def gather_panels(panels):
    """GatherPanels (ParticlesEdgesPanels2.f90:1556-1619): index maps of the
    active (leaf) and passive (divided) panels in tree order."""
    active_map = np.flatnonzero(~panels.hasChildren)
    passive_map = np.flatnonzero(panels.hasChildren)
    return active_map, passive_map


def load_balance(n, num_procs):
    """LoadBalance (PanelVelocityParallel.f90:167-213): contiguous chunks of
    n // P, the last rank also takes the remainder.  Returns 0-based
    half-open (start, end) per rank."""
    chunk = n // num_procs
    starts = [r * chunk for r in range(num_procs)]
    ends = [s + chunk for s in starts]
    ends[-1] = n
    return list(zip(starts, ends))


class PanelVelocity:
    """Drop-in for PanelVelocityParallelModule on one GPU."""

    def __init__(self, device=0, omega=OMEGA_EARTH, smooth=0.05):
        self._ctx = _lib.Context(device)
        self._omega = omega
        self._smooth = smooth
        self._ctx.set_omega(omega)
        self._ctx.set_smooth(smooth)
        self._ready = False          # rk4IsReady
        self._mpi_ready = False
        self._reset_areas = False
        self.totalKE = 0.0
        self.kineticEnergy = np.zeros(0)
        self._active_map = self._passive_map = None

    @property
    def context(self):
        return self._ctx

    # user decisions (:96-122)
    def SetSmooth(self, new_smooth):
        self._smooth = float(new_smooth)
        self._ctx.set_smooth(self._smooth)

    def GetSmooth(self):
        return self._smooth

    def SetOmega(self, new_omega):
        self._omega = float(new_omega)
        self._ctx.set_omega(self._omega)

    def ResetArea(self, on_off=True):
        self._reset_areas = bool(on_off)

    # MPI set-up (:126-229): bookkeeping only on one GPU
    def InitializeMPIRK4(self, particles, panels, proc_rank=0, num_procs=1):
        self._mpi_ready = True
        return self.LoadBalance(particles, panels, num_procs)

    def LoadBalance(self, particles, panels, num_procs=1):
        na = panels.N_Active
        return dict(active=load_balance(na, num_procs),
                    particles=load_balance(particles.N, num_procs),
                    passive=load_balance(panels.N - na, num_procs))

    def FinalizeMPIRK4(self):
        self._mpi_ready = False

    # memory (:236-380)
    def _init_rk4(self, panels):
        self._active_map, self._passive_map = gather_panels(panels)
        a, p = self._active_map, self._passive_map
        self._ctx.load_panels(panels.x[a], panels.relVort[a], panels.area[a], panels.x[p])
        self._ready = True

    def ResetRK4(self):
        if self._ready:
            self._ctx.reset()
        self._ready = False
        self._active_map = self._passive_map = None

    # time stepping (:435-658)
    def BVERK4(self, particles, panels, dt, proc_rank=0, num_procs=1):
        if not self._ready:
            self._init_rk4(panels)
        absv = np.ascontiguousarray(particles.absVort, dtype=np.float64)
        qx, rel, tot, ke = self._ctx.rk4_step(dt, particles.x, absv)
        particles.x[:] = qx
        particles.relVort[:] = rel
        self.totalKE = tot
        self.kineticEnergy = ke
        ax, av, px = self._ctx.store_panels()
        a, p = self._active_map, self._passive_map
        # ScatterPanels: the cached copies overwrite the host tree
        panels.x[a] = ax
        panels.relVort[a] = av
        panels.x[p] = px
        if panels.tracer is not None and panels.tracer.shape[1] >= 2:
            panels.tracer[a, 1] = ke

    def close(self):
        self._ctx.close()


@dataclass
class VorPanels:
    """Voronoi generators in SoA form (VoronoiPanels.f90:25-58)."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    area: np.ndarray
    relVort: np.ndarray
    absVort: np.ndarray = field(default=None)

    @property
    def N(self):
        return self.x.shape[0]


class VoronoiVelocity:
    """Drop-in for VoronoiVelocity{Serial2,Parallel}Module's BVERK4 on one GPU.

    The step ends with NormalizeGrid (VoronoiVelocityParallel.f90:480); the
    re-triangulation the reference does next (DelaunayTri / VoronoiGrid,
    :481-482) stays on the host: pass ``retriangulate`` to run it."""

    def __init__(self, device=0, omega=OMEGA_EARTH):
        self._ctx = _lib.Context(device)
        self._ctx.set_omega(omega)
        self._saved_area = None
        self._use_saved = False
        self.totalKE = 0.0
        self.kineticEnergy = np.zeros(0)

    def UseSavedArea(self, on=True):
        """Freeze the generator areas at their next-step values (Serial2:91-98)."""
        self._use_saved = bool(on)
        self._saved_area = None

    def ResetRK4(self):
        self._saved_area = None

    def BVERK4(self, panels, dt, proc_rank=0, num_procs=1, retriangulate=None):
        if self._use_saved:
            if self._saved_area is None:
                self._saved_area = panels.area.copy()
            area = self._saved_area
        else:
            area = panels.area
        x, y, z, rv, _, ke = self._ctx.voronoi_rk4_step(dt, panels.x, panels.y, panels.z,
                                                        panels.relVort, area)
        # the integral weighs |u|^2 with the current areas, also when the sum
        # uses the saved ones (VoronoiVelocityParallel.f90:358-361)
        self.totalKE = 0.5 * float(np.sum(ke * panels.area))
        panels.x[:], panels.y[:], panels.z[:], panels.relVort[:] = x, y, z, rv
        self.kineticEnergy = ke
        if retriangulate is not None:
            retriangulate(panels)

    def close(self):
        self._ctx.close()
