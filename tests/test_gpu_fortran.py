"""GPU: the Fortran host over the ISO_C_BINDING interface, and the
reference's own drivers (RHWavePanelsMPI.f90, StratospherePanels.f90,
JetPanels.f90 with adaptive refinement, NonDivAdvection.f90; unmodified)
linked against the
Fortran drop-in PanelVelocityParallelModule, against the oracle and the
reference's own output."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, ROOT
from fortran_io import read_dat, read_host_output, write_host_input
from vtk_io import read_frame, rewrite
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic

pytestmark = pytest.mark.gpu
HOST = os.path.join(ROOT, "thesiscode_amd", "bve_host.exe")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "rh4_gpu")
REFERENCE = os.path.join(ROOT, "oracle", "_ref", "rh4_ref")


@pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (no flang)")
def test_fortran_host_matches_oracle(tmp_path):
    g = gridmod.cubed_sphere(4)
    omega = 2 * np.pi
    z = ic.rh4_wave(g.active_x)
    qa = ic.rh4_wave(g.particles_x) + 2 * omega * g.particles_x[:, 2]
    n_steps, dt = 3, 0.01
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_host_input(inp, g.active_x, z, g.active_area, g.passive_x, g.particles_x, qa, n_steps,
                     dt, omega, 0.05)
    r = subprocess.run([HOST, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_host_output(out, g.n_active, g.n_passive, g.n_particles, n_steps)
    st = dict(ax=g.active_x, av=z, px=g.passive_x, qx=g.particles_x)
    kes = []
    for _ in range(n_steps):
        ref = oracle.rk4_step(st["ax"], st["av"], g.active_area, st["px"], st["qx"], qa, dt, omega,
                              0.05)
        st = dict(ax=ref["active_x"], av=ref["active_relvort"], px=ref["passive_x"],
                  qx=ref["particles_x"])
        kes.append(ref["total_ke"])
    disp = np.abs(st["ax"] - g.active_x).max()
    for k, key in (("active_x", "ax"), ("passive_x", "px"), ("particles_x", "qx")):
        assert np.abs(got[k] - st[key]).max() < 1e-11 * disp, k
    assert np.allclose(got["total_ke"], kes, rtol=1e-12, atol=0)
    assert np.abs(got["particles_relvort"] - ref["particles_relvort"]).max() < 1e-12


def run_dropin(tmp_path, namelist, binary=DROPIN, namelist_name="RHWavePanels.namelist",
               precision=None, raw=False):
    """Run a driver in tmp_path on a committed namelist; returns its .dat
    table (its frames are left under tmp_path / vtkOut), or the file's bytes
    with raw.  precision sets BVE_PRECISION for the drop-in modules.  The
    time limit is short: a driver that stops progressing (the reference's
    STRIPACK can loop for ever on some inputs) fails the test."""
    shutil.copy(os.path.join(GOLDEN, namelist), tmp_path / namelist_name)
    (tmp_path / "vtkOut").mkdir()
    env = dict(os.environ)
    env.pop("BVE_PRECISION", None)
    if precision:
        env["BVE_PRECISION"] = precision
    print("driver", os.path.basename(binary), namelist, precision, flush=True)
    r = subprocess.run([binary], cwd=tmp_path, capture_output=True, text=True, timeout=120,
                       env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    dat = list(tmp_path.glob("*.dat"))
    assert len(dat) == 1
    return open(dat[0], "rb").read() if raw else read_dat(dat[0])


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_driver_runs_on_the_dropin(tmp_path):
    """RHWavePanelsMPI (level 4, 12 steps at dt = 0.01) on the GPU module
    reproduces the reference's .dat table -- totalKE, enstrophy and the RH4
    error norms of every step -- to the printed 15 decimals."""
    got = run_dropin(tmp_path, "rh4_driver_L4_noremesh.namelist")
    ref = read_dat(os.path.join(GOLDEN, "rh4_driver_L4_noremesh.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-12, atol=2e-15), np.abs(got - ref).max(axis=0)


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_driver_runs_on_the_dropin_bounded(tmp_path):
    """The same driver with BVE_PRECISION=mixed_bounded (float64 tile
    expansions, float32 remainders, error target 1e-10): its table agrees
    with the reference's to 1e-9 relative over the 12 steps."""
    got = run_dropin(tmp_path, "rh4_driver_L4_noremesh.namelist", precision="mixed_bounded")
    ref = read_dat(os.path.join(GOLDEN, "rh4_driver_L4_noremesh.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-9, atol=1e-15), np.abs(got - ref).max(axis=0)


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_driver_remesh_on_the_dropin(tmp_path):
    """The same with a remesh every 5 steps: the driver calls ResetRK4 and
    the module re-gathers the new grid onto the device.  Rows before the
    first remesh match to the printed precision.  The remesh itself (host
    STRIPACK / SSRFPACK, unchanged reference code) turns rounding-level input
    differences into larger ones -- the reference recompiled with FMA
    contraction moves enstrophy by 1.7e-10 after two remeshes -- and on the
    symmetric cubed sphere it can pick a different but equally valid
    triangulation of cocircular points, so later rows are held to the
    interpolation error scale instead."""
    got = run_dropin(tmp_path, "rh4_driver_L4_remesh5.namelist")
    ref = read_dat(os.path.join(GOLDEN, "rh4_driver_L4_remesh5.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got[:5], ref[:5], rtol=1e-12, atol=2e-15)
    # totalKE and enstrophy after remeshing
    assert np.allclose(got[5:, :2], ref[5:, :2], rtol=1e-8, atol=0)
    # error norms relative to their own size (~1e-3 .. 1e-6)
    assert np.allclose(got[5:, 2:], ref[5:, 2:], rtol=1e-5, atol=1e-12)


@pytest.mark.skipif(not (os.path.exists(DROPIN) and os.path.exists(REFERENCE)),
                    reason="the drivers are built only where the reference sources are")
def test_reference_driver_frames_match_reference_frames(tmp_path):
    """The 13 VTK frames RHWavePanelsMPI writes (one per step: particle
    positions, tracers incl. stage-1 kinetic energy, particle and panel
    vorticity) from the drop-in against those of the reference build run
    here on the host, field by field; and every drop-in frame read back and
    written again by libbve's native writer is byte for byte the file the
    reference's vtkOutput wrote."""
    dirs = {}
    for tag, binary in (("gpu", DROPIN), ("ref", REFERENCE)):
        d = tmp_path / tag
        d.mkdir()
        run_dropin(d, "rh4_driver_L4_noremesh.namelist", binary)
        dirs[tag] = d / "vtkOut"
    names = sorted(os.listdir(dirs["ref"]))
    assert len(names) == 13 and names == sorted(os.listdir(dirs["gpu"]))
    worst = {}
    for name in names:
        got, ref = read_frame(dirs["gpu"] / name), read_frame(dirs["ref"] / name)
        assert got["cells"] == ref["cells"]
        assert np.abs(got["points"] - ref["points"]).max() < 1e-13
        for sec in ("point_data", "cell_data"):
            assert list(got[sec]) == list(ref[sec])
            for key, r in ref[sec].items():
                # relative to the field's size; the error tracer (Tracer3, a
                # small error normalised by max |zeta|) relative to 1
                err = np.abs(got[sec][key] - r).max() / max(np.abs(r).max(), 1.0)
                worst[key] = max(worst.get(key, 0.0), err)
                assert err < 1e-11, (name, key, err)
        out = tmp_path / ("native_" + name)
        rewrite(got, out)
        assert open(out, "rb").read() == open(dirs["gpu"] / name, "rb").read(), name
    print("worst normwise relative difference per field", worst)


STRAT_DROPIN = os.path.join(ROOT, "oracle", "_ref", "strat_gpu")


@pytest.mark.skipif(not os.path.exists(STRAT_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_stratosphere_driver_runs_on_the_dropin(tmp_path):
    """StratospherePanels (level 4, no AMR, 12 StratosphereRK4 steps at
    dt = 0.01 from t = 0) on the GPU module reproduces the reference's .dat
    table -- totalKE and enstrophy of every step -- to the printed 15
    decimals."""
    got = run_dropin(tmp_path, "strat_driver_L4.namelist", STRAT_DROPIN,
                     "StratospherePanels.namelist")
    ref = read_dat(os.path.join(GOLDEN, "strat_driver_L4.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-12, atol=2e-15), np.abs(got - ref).max(axis=0)


JET_DROPIN = os.path.join(ROOT, "oracle", "_ref", "jet_gpu")


@pytest.mark.skipif(not os.path.exists(JET_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_amr_jet_driver_runs_on_the_dropin(tmp_path):
    """JetPanels with adaptive refinement of the initial grid (level 4 plus
    two AMR levels: 3,828 active panels of non-uniform size, 10 steps, no
    remesh) on the GPU module reproduces the reference's .dat table to the
    printed 15 decimals."""
    got = run_dropin(tmp_path, "jet_driver_L4_amr2.namelist", JET_DROPIN, "JetPanels.namelist")
    ref = read_dat(os.path.join(GOLDEN, "jet_driver_L4_amr2.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-12, atol=2e-15), np.abs(got - ref).max(axis=0)


@pytest.mark.skipif(not os.path.exists(JET_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_amr_jet_driver_remesh_on_the_dropin(tmp_path):
    """The same with an adaptive remesh every 5 steps (the grid is re-refined,
    ResetRK4, new panel counts re-gathered onto the device).  Rows before the
    first remesh match to the printed precision.  The remesh is host code
    that can amplify rounding-level differences (see the RH4 remesh test);
    here later rows measured within 6e-13, held to 1e-9."""
    got = run_dropin(tmp_path, "jet_driver_L4_amr2_remesh5.namelist", JET_DROPIN,
                     "JetPanels.namelist")
    ref = read_dat(os.path.join(GOLDEN, "jet_driver_L4_amr2_remesh5.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got[:5], ref[:5], rtol=1e-12, atol=2e-15)
    print("post-remesh max relative difference", np.abs(got[5:] / ref[5:] - 1).max(axis=0))
    assert np.allclose(got[5:], ref[5:], rtol=1e-9, atol=0)


ADV_DROPIN = os.path.join(ROOT, "oracle", "_ref", "adv_gpu")


def read_advection_dat(path):
    """NonDivAdvection's .dat: three error lines, a blank line, then the
    time / totalMass table."""
    lines = open(path).read().splitlines()
    errors = np.array([float(l.split(":")[-1]) for l in lines[:3]])
    table = np.array([[float(v) for v in l.split()] for l in lines[5:] if l.strip()])
    return errors, table


@pytest.mark.skipif(not os.path.exists(ADV_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_advection_driver_runs_on_the_dropin(tmp_path):
    """NonDivAdvection (Gaussian hills, level 3, AdvectionRK4 over one full
    period, 100 steps, no remesh) on the GPU module: the tracer errors at
    t = 5 (~2e-5) agree with the reference's to 1e-8 relative (positions
    differ at the 1e-15 level), the mass history exactly."""
    shutil.copy(os.path.join(GOLDEN, "adv_driver_L3.namelist"),
                tmp_path / "AdvectGaussHills.namelist")
    (tmp_path / "vtkOut").mkdir()
    r = subprocess.run([ADV_DROPIN], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-2000:]
    dat = list(tmp_path.glob("*.dat"))
    assert len(dat) == 1
    got_err, got_tab = read_advection_dat(dat[0])
    ref_err, ref_tab = read_advection_dat(os.path.join(GOLDEN, "adv_driver_L3.dat"))
    assert np.allclose(got_err, ref_err, rtol=1e-8, atol=0), (got_err, ref_err)
    assert np.array_equal(got_tab, ref_tab)


VOR_DROPIN = os.path.join(ROOT, "oracle", "_ref", "vor_gpu")


@pytest.mark.skipif(not os.path.exists(VOR_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_voronoi_driver_runs_on_the_dropin(tmp_path):
    """GaussVortVoronoi (icosahedral Voronoi grid, 642 generators, 20 steps
    at dt = 0.01 summed with the saved initial areas, the reference's
    STRIPACK re-triangulation after every step) linked against the drop-in
    VoronoiVelocityParallelModule, in float64 (BVE_PRECISION=fp64),
    reproduces the reference's .dat table -- totalKE and enstrophy of every
    step -- to the printed 15 decimals."""
    got = run_dropin(tmp_path, "vor_driver_L3.namelist", VOR_DROPIN, "GaussVortVoronoi.namelist",
                     precision="fp64")
    ref = read_dat(os.path.join(GOLDEN, "vor_driver_L3.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-12, atol=2e-15), np.abs(got - ref).max(axis=0)


VORS_DROPIN = os.path.join(ROOT, "oracle", "_ref", "vors_gpu")


@pytest.mark.skipif(not os.path.exists(VORS_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_serial_voronoi_driver_runs_on_the_dropin(tmp_path):
    """RHWaveVoronoi (the serial driver: 642 generators, 20 steps at dt = 0.01,
    saved areas) linked against the drop-in VoronoiVelocitySerialModule, in
    the Voronoi modules' default reference-order arithmetic, writes the same
    .dat file, byte for byte, as the driver linked against the reference's
    VoronoiVelocitySerial2.f90: the positions handed to STRIPACK are the
    reference's bit for bit.  (With float64 rounding instead, the reference's
    STRIPACK loops for ever after step 9 of this trajectory: shown on the CPU
    with the GPU kernel's arithmetic form.)"""
    got = run_dropin(tmp_path, "vors_driver_L3.namelist", VORS_DROPIN, "RHWaveVoronoi.namelist",
                     raw=True)
    assert got == open(os.path.join(GOLDEN, "vors_driver_L3.dat"), "rb").read()


RH4S_DROPIN = os.path.join(ROOT, "oracle", "_ref", "rh4s_gpu")


def check_serial_panel_table(got, ref, exact):
    """The serial RH4 driver (RHWavePanels.f90) records the l1 / l2 / linf
    vorticity errors of every step but, unlike the MPI driver, never stores
    totalKE or the enstrophy after step 0 (its arrays keep their zeros):
    columns 3-5 and row 0 are compared with the MPI driver's table."""
    assert got.shape == ref.shape
    assert np.array_equal(got[1:, :2], np.zeros_like(got[1:, :2]))
    assert np.array_equal(got[0], ref[0])
    if exact:
        assert np.array_equal(got[:, 2:], ref[:, 2:]), np.abs(got - ref)[:, 2:].max(axis=0)
    else:
        assert np.allclose(got[:, 2:], ref[:, 2:], rtol=1e-12, atol=2e-15), \
            np.abs(got - ref)[:, 2:].max(axis=0)


@pytest.mark.skipif(not os.path.exists(RH4S_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_serial_panel_driver_runs_on_the_dropin(tmp_path):
    """RHWavePanels (the serial panel driver, which the reference's own
    PanelVelocitySerial.f90 cannot be compiled for here) linked against the
    drop-in PanelVelocitySerialModule in float64, on the namelist of the MPI
    driver's golden run (level 4, 12 steps, no remesh): the RH4 error norms
    of every step equal those the reference's RHWavePanelsMPI wrote with the
    reference's PanelVelocityParallelModule, to the printed 15 decimals."""
    got = run_dropin(tmp_path, "rh4_driver_L4_noremesh.namelist", RH4S_DROPIN, precision="fp64")
    check_serial_panel_table(got, read_dat(os.path.join(GOLDEN, "rh4_driver_L4_noremesh.dat")),
                             exact=False)


@pytest.mark.skipif(not os.path.exists(RH4S_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_serial_panel_driver_bitwise_in_reference_mode(tmp_path):
    """As above with BVE_PRECISION=reference: every printed digit of the
    error norms equals the reference MPI driver's."""
    got = run_dropin(tmp_path, "rh4_driver_L4_noremesh.namelist", RH4S_DROPIN,
                     precision="reference")
    check_serial_panel_table(got, read_dat(os.path.join(GOLDEN, "rh4_driver_L4_noremesh.dat")),
                             exact=True)


@pytest.mark.skipif(not os.path.exists(VOR_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_voronoi_driver_bitwise_in_reference_mode(tmp_path):
    """GaussVortVoronoi on the drop-in in its default reference-order
    arithmetic: the .dat file equals the reference run's byte for byte."""
    got = run_dropin(tmp_path, "vor_driver_L3.namelist", VOR_DROPIN, "GaussVortVoronoi.namelist",
                     raw=True)
    assert got == open(os.path.join(GOLDEN, "vor_driver_L3.dat"), "rb").read()


@pytest.mark.skipif(not os.path.exists(DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
@pytest.mark.parametrize("name", ["rh4_driver_L4_noremesh", "rh4_driver_L4_remesh5"])
def test_reference_driver_bitwise_in_reference_mode(tmp_path, name):
    """RHWavePanelsMPI on the drop-in with BVE_PRECISION=reference writes the
    reference run's .dat file byte for byte -- also across the remesh every
    5 steps, where the float64 mode can only follow the reference to the
    interpolation-error scale."""
    got = run_dropin(tmp_path, name + ".namelist", precision="reference", raw=True)
    assert got == open(os.path.join(GOLDEN, name + ".dat"), "rb").read()


@pytest.mark.skipif(not os.path.exists(JET_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
@pytest.mark.parametrize("name", ["jet_driver_L4_amr2", "jet_driver_L4_amr2_remesh5"])
def test_reference_amr_jet_driver_bitwise_in_reference_mode(tmp_path, name):
    """JetPanels (adaptively refined grid, with and without adaptive remesh)
    on the drop-in with BVE_PRECISION=reference: the reference run's .dat
    file byte for byte."""
    got = run_dropin(tmp_path, name + ".namelist", JET_DROPIN, "JetPanels.namelist",
                     precision="reference", raw=True)
    assert got == open(os.path.join(GOLDEN, name + ".dat"), "rb").read()


GAUSS_DROPIN = os.path.join(ROOT, "oracle", "_ref", "gauss_gpu")


@pytest.mark.skipif(not os.path.exists(GAUSS_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
def test_reference_gauss_vortex_driver_runs_on_the_dropin(tmp_path):
    """GaussVort (the Gaussian-vortex panel driver on icosahedral triangles,
    level 4: 5,120 active panels, 10 steps at dt = 0.01, no remesh) on the
    GPU module reproduces the reference's .dat table -- totalKE and
    enstrophy of every step -- to the printed 15 decimals."""
    got = run_dropin(tmp_path, "gauss_driver_tri_L4.namelist", GAUSS_DROPIN,
                     "GaussVortPanels.namelist", precision="fp64")
    ref = read_dat(os.path.join(GOLDEN, "gauss_driver_tri_L4.dat"))
    assert got.shape == ref.shape
    assert np.allclose(got, ref, rtol=1e-12, atol=2e-15), np.abs(got - ref).max(axis=0)


@pytest.mark.skipif(not os.path.exists(GAUSS_DROPIN),
                    reason="the drop-in driver is built only where the reference sources are")
@pytest.mark.parametrize("name", ["gauss_driver_tri_L4", "gauss_driver_tri_L3_remesh5"])
def test_reference_gauss_vortex_driver_bitwise_in_reference_mode(tmp_path, name):
    """GaussVort on triangles, with and without a remesh every 5 steps, on
    the drop-in with BVE_PRECISION=reference: the reference run's .dat file
    byte for byte."""
    got = run_dropin(tmp_path, name + ".namelist", GAUSS_DROPIN, "GaussVortPanels.namelist",
                     precision="reference", raw=True)
    assert got == open(os.path.join(GOLDEN, name + ".dat"), "rb").read()
