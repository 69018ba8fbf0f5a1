"""Uniform cubed-sphere panel grid on the unit sphere (host side).

Builds the grid the reference's ``InitCubedSphereGrid`` / ``DivideQuadPanel``
produce (ParticlesEdgesPanels2.f90:648-757, :1093-1535) for a uniform
refinement level L, without the edge tree the reference keeps for AMR:

* the root cube: 8 corners at (+-a, +-a, +-a), a = 1/sqrt(3), and 6 face
  panels centred on +-e_k with the reference's corner order (:662-739);
* a panel with corners (v1, v2, v3, v4) and centre c divides into
  [v1, m12, c, m41], [m12, v2, m23, c], [c, m23, v3, m34], [m41, c, m34, v4]
  with m_ab = SphereMidpoint(a, b) (SphereGeom.f90:91) and child centres
  SphereQuadCenter of the child's corners (SphereGeom.f90:114);
* panel areas are the four L'Huilier triangles (corner, centre, next corner)
  of QuadPanelArea (ParticlesEdgesPanels2.f90:1804-1895) with the
  SphereTriArea / SphereDistance formulas (SphereGeom.f90:50-68, :126-145);
* divided panels get area 0 and become "passive"; the panel order is the
  reference's breadth-first tree order, so active panels come out in the
  order GatherPanels (:1556-1619) gives them.

Particles (panel corners) are the distinct corners of the leaf panels, in
first-appearance order over the leaves; the reference orders them by edge
division instead, so they match its particles as a set, not by index.
Counts: N_active = 6*4^L, N_particles = 6*4^L + 2, N_passive = 2*(4^L - 1).
"""
from dataclasses import dataclass

import numpy as np

_A = 1.0 / np.sqrt(3.0)
# root cube corners, ParticlesEdgesPanels2.f90:662-669
_CORNERS = np.array([[_A, -_A, _A], [_A, -_A, -_A], [_A, _A, -_A], [_A, _A, _A],
                     [-_A, _A, -_A], [-_A, _A, _A], [-_A, -_A, -_A], [-_A, -_A, _A]])
# face corner indices (0-based) and centres, :727-739
_FACE_VERTS = np.array([[0, 1, 2, 3], [3, 2, 4, 5], [5, 4, 6, 7], [7, 6, 1, 0],
                        [7, 0, 3, 5], [1, 6, 4, 2]])
_FACE_CENTRES = np.array([[1.0, 0, 0], [0, 1.0, 0], [-1.0, 0, 0], [0, -1.0, 0],
                          [0, 0, 1.0], [0, 0, -1.0]])


def _normalize(v):
    # x / sqrt(sum(x*x)), summed left to right as the reference does
    r = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])
    return v / r[..., None]


def sphere_midpoint(a, b):
    return _normalize((a + b) / 2.0)


def sphere_quad_center(a, b, c, d):
    return _normalize((((a + b) + c) + d) / 4.0)


def sphere_distance(a, b):
    cp0 = a[..., 1] * b[..., 2] - b[..., 1] * a[..., 2]
    cp1 = b[..., 0] * a[..., 2] - a[..., 0] * b[..., 2]
    cp2 = a[..., 0] * b[..., 1] - b[..., 0] * a[..., 1]
    cn = np.sqrt((cp0 * cp0 + cp1 * cp1) + cp2 * cp2)
    dp = (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]
    return np.arctan2(cn, dp)


def sphere_tri_area(a, b, c):
    s1 = sphere_distance(a, b)
    s2 = sphere_distance(b, c)
    s3 = sphere_distance(c, a)
    hp = ((s1 + s2) + s3) / 2.0
    zz = ((np.tan(hp / 2.0) * np.tan((hp - s1) / 2.0)) * np.tan((hp - s2) / 2.0)) * np.tan((hp - s3) / 2.0)
    return 4.0 * np.arctan2(np.sqrt(zz), 1.0)


def quad_area(verts, centre):
    """Sum of the four triangles (v_k, centre, v_{k+1}); verts (n, 4, 3)."""
    area = np.zeros(verts.shape[0])
    for k in range(4):
        area = area + sphere_tri_area(verts[:, k], centre, verts[:, (k + 1) % 4])
    return area


@dataclass
class CubedSphere:
    """Uniform cubed sphere: the arrays the velocity path consumes."""
    level: int
    active_x: np.ndarray       # (N_active, 3) leaf panel centres
    active_area: np.ndarray    # (N_active,)
    passive_x: np.ndarray      # (N_passive, 3) divided (parent) panel centres
    particles_x: np.ndarray    # (N_particles, 3) panel corners
    active_verts: np.ndarray   # (N_active, 4) indices into particles_x

    @property
    def n_active(self):
        return self.active_x.shape[0]

    @property
    def n_passive(self):
        return self.passive_x.shape[0]

    @property
    def n_particles(self):
        return self.particles_x.shape[0]


def cubed_sphere(level):
    """Build the uniform cubed sphere at refinement ``level`` (0 = the cube)."""
    if level < 0 or level > 10:
        raise ValueError("level must be in [0, 10]")
    verts = _CORNERS[_FACE_VERTS]          # (6, 4, 3)
    centres = _FACE_CENTRES.copy()
    passive = []
    for _ in range(level):
        passive.append(centres)
        v1, v2, v3, v4 = verts[:, 0], verts[:, 1], verts[:, 2], verts[:, 3]
        c = centres
        m12 = sphere_midpoint(v1, v2)
        m23 = sphere_midpoint(v2, v3)
        m34 = sphere_midpoint(v3, v4)
        m41 = sphere_midpoint(v4, v1)
        kids = np.stack([np.stack([v1, m12, c, m41], 1),
                         np.stack([m12, v2, m23, c], 1),
                         np.stack([c, m23, v3, m34], 1),
                         np.stack([m41, c, m34, v4], 1)], 1)   # (n, 4 children, 4, 3)
        verts = kids.reshape(-1, 4, 3)
        centres = sphere_quad_center(verts[:, 0], verts[:, 1], verts[:, 2], verts[:, 3])
    area = quad_area(verts, centres)
    passive_x = np.concatenate(passive) if passive else np.zeros((0, 3))
    # distinct corners, exact (shared-edge midpoints are bitwise equal)
    flat = np.ascontiguousarray(verts.reshape(-1, 3))
    keys = flat.view(np.dtype((np.void, 24))).ravel()
    _, first, inv = np.unique(keys, return_index=True, return_inverse=True)
    order = np.argsort(first)
    rank = np.empty_like(order)
    rank[order] = np.arange(order.size)
    particles = flat[first[order]]
    active_verts = rank[inv.ravel()].reshape(-1, 4)
    return CubedSphere(level, np.ascontiguousarray(centres), area, passive_x,
                       np.ascontiguousarray(particles), active_verts)


def counts(level):
    """(N_active, N_particles, N_passive) of the uniform cubed sphere."""
    n = 6 * 4 ** level
    return n, n + 2, 2 * (4 ** level - 1)
