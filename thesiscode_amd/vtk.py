"""VTK frames in the reference's format, written by libbve's native writer.

``vtkOutput(particles, panels, filename)`` mirrors the reference's
``vtkOutput`` / ``vtkOutputUnif`` (ParticlesEdgesPanels2.f90:133-137,
:2116-2260): one legacy-VTK ASCII POLYDATA file with the particle positions
as POINTS, the undivided panels as POLYGONS, per-particle tracers / relVort
/ absVort as POINT_DATA and per-undivided-panel tracers / relVort / absVort
as CELL_DATA, with the reference's F24.18 and I8 fields, byte for byte.
The formatting runs in libbve (bve_vtk_write_frame); there is no Python
fallback.
"""
import numpy as np

from . import _lib

TRI_PANEL, QUAD_PANEL, VORONOI_PANEL = 3, 4, 10   # ParticlesEdgesPanels2.f90:36-38


def _col(a, n, what):
    """float64 array of n rows in column-major order (Fortran tracer(N, k))."""
    a = np.asfortranarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != n:
        raise ValueError(f"{what}: expected shape ({n}, n_tracer), got {a.shape}")
    return a


def _vec(a, n, what):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != (n,):
        raise ValueError(f"{what}: expected shape ({n},), got {a.shape}")
    return a


def write_frame(filename, particles_x, panel_kind, vertices, has_children=None, n_active=None,
                particles_relvort=None, particles_absvort=None, particles_tracer=None,
                panels_tracer=None, panels_relvort=None, panels_absvort=None):
    """Write one frame.

    particles_x (N, 3); vertices (n_panels, panel_kind) 0-based particle
    indices, -1 for an unused Voronoi slot; has_children (n_panels,) bool or
    None; n_active defaults to the number of undivided panels;
    particles_tracer (N, n_tracer) and panels_tracer (n_panels, n_tracer) or
    None; the vorticity arrays or None when the problem carries none.
    """
    L = _lib.load()
    x = np.ascontiguousarray(particles_x, dtype=np.float64)
    if x.ndim != 2 or x.shape[1] != 3:
        raise ValueError(f"particles_x: expected shape (N, 3), got {x.shape}")
    n = x.shape[0]
    if panel_kind not in (TRI_PANEL, QUAD_PANEL, VORONOI_PANEL):
        raise ValueError("panel_kind must be 3, 4 or 10")
    v = np.asarray(vertices)
    if v.ndim != 2 or v.shape[1] != panel_kind:
        raise ValueError(f"vertices: expected shape (n_panels, {panel_kind}), got {v.shape}")
    n_panels = v.shape[0]
    v1 = np.ascontiguousarray(v, dtype=np.int32) + np.int32(1)     # the reference's 1-based indices
    hc = None
    if has_children is not None:
        hc = np.ascontiguousarray(np.asarray(has_children, dtype=bool).astype(np.int32))
        if hc.shape != (n_panels,):
            raise ValueError(f"has_children: expected shape ({n_panels},), got {hc.shape}")
    if n_active is None:
        n_active = n_panels - (0 if hc is None else int(hc.sum()))
    pt = None if particles_tracer is None else _col(particles_tracer, n, "particles_tracer")
    n_tracer = 0 if pt is None else pt.shape[1]
    qt = None
    if panels_tracer is not None and n_tracer > 0:
        qt = _col(panels_tracer, n_panels, "panels_tracer")
        if qt.shape[1] < n_tracer:
            raise ValueError("panels_tracer has fewer tracers than particles_tracer")
    prv = _vec(particles_relvort, n, "particles_relvort")
    pav = _vec(particles_absvort, n, "particles_absvort")
    qrv = _vec(panels_relvort, n_panels, "panels_relvort")
    qav = _vec(panels_absvort, n_panels, "panels_absvort")
    P = _lib._ptr
    rc = L.bve_vtk_write_frame(str(filename).encode(), n, n, P(x), P(prv), P(pav), n_tracer, P(pt),
                               panel_kind, n_panels, int(n_active), n_panels, P(v1), P(hc), P(qt),
                               P(qrv), P(qav))
    if rc != _lib.BVE_OK:
        raise _lib.BVEError(f"bve_vtk_write_frame failed ({rc}): {L.bve_last_error(None).decode()}")


def vtkOutput(particles, panels, filename):
    """The reference's vtkOutput(aParticles, aPanels, filename) on the
    host-side Particles / Panels (thesiscode_amd.panel_velocity); the panels
    need their corner lists (``panels.vertices``, 0-based) and kind."""
    if panels.vertices is None:
        raise ValueError("panels.vertices is required to write cells")
    write_frame(filename, particles.x, panels.panelKind, panels.vertices,
                has_children=panels.hasChildren, n_active=panels.N_Active,
                particles_relvort=particles.relVort, particles_absvort=particles.absVort,
                particles_tracer=particles.tracer, panels_tracer=panels.tracer,
                panels_relvort=panels.relVort, panels_absvort=panels.absVort)
