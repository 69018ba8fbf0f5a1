"""VTK frame output (vtkOutputUnif, ParticlesEdgesPanels2.f90:2116-2260)
through libbve's native writer: byte for byte against frames the reference's
own routine wrote, F24.18 rounding against Python's correctly rounded
decimal conversion, the host-side Panels mirror, errors.  Host code only."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from thesiscode_amd import _lib, vtk
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic
from thesiscode_amd import panel_velocity as pv
from vtk_io import read_frame

NATIVE_CHECK = os.path.join(ROOT, "oracle", "_ref", "vtk_native_check")


def vtk_cases():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)["vtk_cases"]


def write_golden_case(name, out):
    with np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    vtk.write_frame(out, g["particles_x"], int(g["panel_kind"][0]), g["vertices"] - 1,
                    has_children=g["has_children"].astype(bool), n_active=int(g["n_active"][0]),
                    particles_relvort=g.get("particles_relvort"),
                    particles_absvort=g.get("particles_absvort"),
                    particles_tracer=g["particles_tracer"] if g["particles_tracer"].shape[1] else None,
                    panels_tracer=g["panels_tracer"], panels_relvort=g.get("panels_relvort"),
                    panels_absvort=g.get("panels_absvort"))
    return g


@pytest.mark.parametrize("name", sorted(vtk_cases()))
def test_frame_is_byte_identical_to_reference(tmp_path, name):
    """Cases: a cubed sphere with locally divided panels after a BVERK4
    step, icosahedral triangles without vorticity, and the Voronoi kind with
    wrapped corner lists and the F24.18 edge values (signed zeros, ties,
    sub-normals, fields too wide, NaN, infinities)."""
    out = tmp_path / "frame.vtk"
    write_golden_case(name, out)
    ref = open(os.path.join(GOLDEN, name + ".vtk"), "rb").read()
    got = open(out, "rb").read()
    if got != ref:
        for k, (a, b) in enumerate(zip(got.split(b"\n"), ref.split(b"\n"))):
            assert a == b, f"line {k + 1}: {a!r} != {b!r}"
    assert got == ref


def f24_18(v):
    """Fw.d with w = 24, d = 18 as the Fortran runtime writes it, from
    Python's correctly rounded (round-half-even) decimal conversion."""
    if np.isnan(v):
        s = "NaN"
    elif np.isinf(v):
        s = "Inf" if v > 0 else "-Inf"
    else:
        s = "%.18f" % abs(v)
        if np.signbit(v):
            s = "-" + s
    return "*" * 24 if len(s) > 24 else s.rjust(24)


def test_rounding_matches_correctly_rounded_decimal(tmp_path):
    rng = np.random.default_rng(7)
    vals = [rng.uniform(-1, 1, 3000) * 10.0 ** rng.integers(-21, 6, 3000)]
    # exact ties at the 18th decimal: k * 2^-j with j <= 60 has at most 60 decimals
    k = rng.integers(1, 2 ** 20, 2000)
    j = rng.integers(19, 61, 2000)
    vals.append(k * 2.0 ** -j.astype(float))
    vals.append(np.array([5e-19, 1.5e-18, 2.5e-18, 0.5, 99999.5, 99999.99999999999, 1e5,
                          -9999.5, -1e4, 131071.9, 131072.0, 5e-324, -0.0, 0.0, np.nan, np.inf,
                          -np.inf, 1 / 3]))
    v = np.concatenate(vals)
    v = v[: 3 * (len(v) // 3)].reshape(-1, 3)
    out = tmp_path / "f.vtk"
    vtk.write_frame(out, v, vtk.QUAD_PANEL, np.zeros((0, 4), np.int64))
    rows = open(out).read().split("\n")[5:5 + len(v)]
    expect = ["".join(f24_18(x) for x in r) for r in v]
    bad = [(a, b) for a, b in zip(rows, expect) if a != b]
    assert not bad, bad[:5]


def test_host_panels_mirror_frame(tmp_path):
    """vtkOutput on the host-side Particles / Panels of a uniform cubed
    sphere: counts, cells over the undivided panels' corners, and every
    value read back to within the F24.18 quantum."""
    g = gridmod.cubed_sphere(3)
    z = ic.rh4_wave(g.active_x)
    panels = pv.panels_from_grid(g, z)
    panels.tracer[:, 0] = panels.x[:, 2]
    particles = pv.Particles(x=g.particles_x, relVort=ic.rh4_wave(g.particles_x),
                             absVort=ic.rh4_wave(g.particles_x) + 1.0,
                             tracer=np.column_stack([g.particles_x[:, 0], -g.particles_x[:, 1],
                                                     g.particles_x[:, 2]]))
    out = tmp_path / "mirror.vtk"
    vtk.vtkOutput(particles, panels, out)
    fr = read_frame(out)
    assert fr["n_active"] == g.n_active and fr["points"].shape == (g.n_particles, 3)
    assert np.abs(fr["points"] - g.particles_x).max() <= 5e-19
    cells = np.array(fr["cells"])
    assert cells.shape == (g.n_active, 5) and (cells[:, 0] == 4).all()
    assert np.array_equal(cells[:, 1:], g.active_verts)
    assert list(fr["point_data"]) == ["Tracer1", "Tracer2", "Tracer3", "relVort", "absVort"]
    assert list(fr["cell_data"]) == ["TracerPanel1", "TracerPanel2", "TracerPanel3",
                                     "relVortPanel", "absVortPanel"]
    leaves = ~panels.hasChildren
    assert np.abs(fr["cell_data"]["relVortPanel"] - panels.relVort[leaves]).max() <= 5e-19 * 10
    assert np.abs(fr["cell_data"]["TracerPanel1"] - panels.x[leaves, 2]).max() <= 5e-19
    assert np.abs(fr["point_data"]["Tracer2"] + g.particles_x[:, 1]).max() <= 5e-19


def test_writer_errors(tmp_path):
    L = _lib.load()
    x = np.zeros((2, 3))
    with pytest.raises(ValueError):
        vtk.write_frame(tmp_path / "a.vtk", x, 5, np.zeros((1, 5), np.int64))
    rc = L.bve_vtk_write_frame(str(tmp_path / "b.vtk").encode(), 2, 2, _lib._ptr(x), None, None,
                               0, None, 5, 0, 0, 0, None, None, None, None, None)
    assert rc == 1 and b"panel_kind" in L.bve_last_error(None)
    with pytest.raises(_lib.BVEError, match="cannot open"):
        vtk.write_frame(tmp_path / "missing" / "c.vtk", x, 4, np.zeros((0, 4), np.int64))
    rc = L.bve_vtk_write_frame(None, 0, 0, None, None, None, 0, None, 4, 0, 0, 0, None, None,
                               None, None, None)
    assert rc == 1


@pytest.mark.skipif(not os.path.exists(NATIVE_CHECK),
                    reason="built only where the reference sources are")
@pytest.mark.parametrize("kind", [3, 4])
def test_fortran_native_writer_matches_reference_writer(tmp_path, kind):
    """vtkOutputNative (Fortran, over the C ABI) and the reference's
    vtkOutput on the same reference-built grid write identical files."""
    r = subprocess.run([NATIVE_CHECK, str(kind), "3", "ref.vtk", "native.vtk"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert open(tmp_path / "ref.vtk", "rb").read() == open(tmp_path / "native.vtk", "rb").read()
