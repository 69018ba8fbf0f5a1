"""Write the golden vectors in this folder from the reference's own Fortran.

Runs ``oracle/_ref/make_golden`` (built by ``make -f oracle/Makefile.ref``
from the reference's unmodified sources; see tests/golden/make_golden.f90)
for a fixed list of cases and stores each case's inputs and outputs as one
compressed ``.npz`` (float64 / int32 arrays only, loadable with
``allow_pickle=False``) plus ``manifest.json``.

It also runs the reference's own drivers (RHWavePanelsMPI.f90,
StratospherePanels.f90, JetPanels.f90, NonDivAdvection.f90, GaussVort.f90;
unmodified,
built by the same makefile's `drivers` target) on the committed namelists
and keeps their .dat tables.  ``make_golden.py voronoi`` redoes only the
Voronoi cases (make_voronoi_golden.f90 and the two Voronoi drivers).

Run from the repository root, in a container that has the reference:

    make -f oracle/Makefile.ref all drivers && python tests/golden/make_golden.py

It also writes VTK frames with the reference's vtkOutput (make_vtk_golden.f90)
together with the arrays they were written from.
"""
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
HARNESS = os.path.join(ROOT, "oracle", "_ref", "make_golden")
VTK_HARNESS = os.path.join(ROOT, "oracle", "_ref", "make_vtk_golden")

# name: (level, ic, nsteps, dt, smooth, description)
CASES = {
    "rh4ref_L4_1step": (4, 0, 1, 0.01, 0.05,
                        "cubed sphere L4, reference RH4 wave (InitRH4Wave), one RK4 step"),
    "rh4_L4_12steps": (4, 1, 12, 0.01, 0.05,
                       "cubed sphere L4, RH4 with omega=K=7.848e-6*86400, 12 RK4 steps dt=0.01"),
    "gauss_L3_2steps_smooth010": (3, 2, 2, 0.01, 0.10,
                                  "cubed sphere L3, reference Gaussian vortex, smooth=0.10, 2 steps"),
    "rh4_L5_1step": (5, 1, 1, 0.01, 0.05,
                     "cubed sphere L5, RH4 with omega=K=7.848e-6*86400, one RK4 step"),
}

# forced (StratosphereRK4) cases: name: (level, nsteps, dt, smooth, t0, description)
FORCED_CASES = {
    "strat_L4_3steps_t1": (4, 3, 0.01, 0.05, 1.0,
                           "cubed sphere L4, InitStratosphere, 3 StratosphereRK4 steps from t=1 "
                           "(forcing ramping up)"),
    "strat_L3_2steps_t399": (3, 2, 0.01, 0.05, 3.99,
                             "cubed sphere L3, InitStratosphere, 2 StratosphereRK4 steps from "
                             "t=3.99 (stage times cross the end of the ramp at t=4)"),
    "strat_L3_2steps_t5": (3, 2, 0.05, 0.10, 5.0,
                           "cubed sphere L3, InitStratosphere, 2 StratosphereRK4 steps dt=0.05 "
                           "from t=5 (full forcing), smooth=0.10"),
}


# advection (AdvectionRK4 ic 4, DivergentAdvectionRK4 ic 5) cases:
# name: (level, ic, nsteps, dt, t0, description)
ADVECTION_CASES = {
    "adv_nondiv_L4_4steps_t1": (4, 4, 4, 0.05, 1.0,
                                "cubed sphere L4, AdvectionRK4 (NonDivergentWind), 4 steps dt=0.05 "
                                "from t=1"),
    "adv_div_L4_4steps_t1": (4, 5, 4, 0.05, 1.0,
                             "cubed sphere L4, DivergentAdvectionRK4 (DivergentWind), 4 steps "
                             "dt=0.05 from t=1"),
    "adv_nondiv_L3_3steps_t24": (3, 4, 3, 0.05, 2.4,
                                 "cubed sphere L3, AdvectionRK4 across t=2.5 (wind reduces to a "
                                 "solid-body rotation there)"),
}

# reference drivers run unmodified on a committed namelist: name: (binary in
# oracle/_ref, the namelist file name the driver opens); the namelist is
# tests/golden/<name>.namelist and the driver's .dat becomes <name>.dat
DRIVER_CASES = {
    "rh4_driver_L4_noremesh": ("rh4_ref", "RHWavePanels.namelist"),
    "rh4_driver_L4_remesh5": ("rh4_ref", "RHWavePanels.namelist"),
    "strat_driver_L4": ("strat_ref", "StratospherePanels.namelist"),
    "jet_driver_L4_amr2": ("jet_ref", "JetPanels.namelist"),
    "jet_driver_L4_amr2_remesh5": ("jet_ref", "JetPanels.namelist"),
    "adv_driver_L3": ("adv_ref", "AdvectGaussHills.namelist"),
    "gauss_driver_tri_L4": ("gauss_ref", "GaussVortPanels.namelist"),
    "gauss_driver_tri_L3_remesh5": ("gauss_ref", "GaussVortPanels.namelist"),
}
# the reference's Voronoi Gaussian-vortex driver (GaussVortVoronoi.f90, MPI
# module) and its serial RH4 Voronoi driver (RHWaveVoronoi.f90, serial module)
VORONOI_DRIVER_CASES = {
    "vor_driver_L3": ("vor_ref", "GaussVortVoronoi.namelist"),
    "vors_driver_L3": ("vors_ref", "RHWaveVoronoi.namelist"),
}


# Voronoi BVERK4 (VoronoiVelocityParallel.f90:304-483) on icosahedral Voronoi
# grids (tests/golden/make_voronoi_golden.f90):
# name: (initNest, ic, nsteps, dt, saved_area, description)
VORONOI_CASES = {
    "vor_rh4_L3_1step": (3, 0, 1, 0.01, 0,
                         "icosahedral Voronoi grid, 642 generators, reference RH4 wave, one step"),
    "vor_rh4_L4_2steps": (4, 0, 2, 0.01, 0,
                          "icosahedral Voronoi grid, 2562 generators, reference RH4 wave, two "
                          "steps, areas re-tessellated after each"),
    "vor_rh4_L3_3steps_saved": (3, 0, 3, 0.01, 1,
                                "642 generators, reference RH4 wave, three steps summed with the "
                                "initial areas (UseSavedArea)"),
    "vor_gauss_L3_2steps": (3, 1, 2, 0.02, 0,
                            "642 generators, reference Gaussian vortex, two steps dt=0.02"),
}
VORONOI_HARNESS = os.path.join(ROOT, "oracle", "_ref", "make_voronoi_golden")


def read_voronoi_dump(path):
    """Parse the stream file written by make_voronoi_golden.f90."""
    raw = open(path, "rb").read()
    n, nsteps, saved = (int(v) for v in np.frombuffer(raw, "<i4", 3, 0))
    dt, omega = np.frombuffer(raw, "<f8", 2, 12)
    off = 28

    def take(count):
        nonlocal off
        a = np.frombuffer(raw, "<f8", count, off).copy()
        off += 8 * count
        return a

    d = dict(n=n, nsteps=nsteps, saved=saved, dt=float(dt), omega=float(omega))
    arrays = {}
    for s in range(nsteps + 1):
        for key in ("x", "y", "z", "relvort", "area"):
            arrays[f"s{s}_{key}"] = take(n)
    for s in range(1, nsteps + 1):
        arrays[f"ke_step{s}"] = take(n)
        arrays[f"total_ke_step{s}"] = take(1)
    assert off == len(raw), (off, len(raw))
    return d, arrays


def run_voronoi_case(tmp, name, nest, ic, nsteps, dt, saved):
    out = os.path.join(tmp, name + ".bin")
    # the reference's STRIPACK can loop for ever on a degenerate point set,
    # hence the time limit
    subprocess.run([VORONOI_HARNESS, str(nest), str(ic), str(nsteps), repr(dt), str(saved), out],
                   check=True, cwd=tmp, stdout=subprocess.DEVNULL, timeout=120)
    d, arrays = read_voronoi_dump(out)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print(name, d["n"])
    return d


# VTK frames written by the reference's vtkOutput (tests/golden/make_vtk_golden.f90):
# name: (harness case, description); stored as <name>.vtk plus <name>.npz
VTK_CASES = {
    "vtk_quad_L2_amr": (1, "cubed sphere L2 with 8 panels divided once more, RH4 wave after one "
                           "BVERK4 step, 3 tracers"),
    "vtk_tri_L2_advection": (2, "icosahedral triangles L2, advection problem (no vorticity), "
                                "2 tracers"),
    "vtk_voronoi_special": (3, "Voronoi panel kind, synthetic: wrapped list-directed corner lists; "
                               "signed zeros, ties, sub-normals, too-wide values, NaN, Inf"),
}


def read_vtk_dump(path):
    """Parse the stream file written by make_vtk_golden.f90."""
    raw = open(path, "rb").read()
    kind, n, npan, nact, nt, has_vort = (int(v) for v in np.frombuffer(raw, "<i4", 6, 0))
    off = 24

    def take(count, dtype="<f8"):
        nonlocal off
        a = np.frombuffer(raw, dtype, count, off).copy()
        off += a.itemsize * count
        return a

    d = dict(panel_kind=np.array([kind], np.int32), n_active=np.array([nact], np.int32))
    d["particles_x"] = take(3 * n).reshape(n, 3)
    if has_vort:
        d["particles_relvort"] = take(n)
        d["particles_absvort"] = take(n)
    d["particles_tracer"] = take(n * nt).reshape(nt, n).T.copy()
    d["vertices"] = take(kind * npan, "<i4").reshape(npan, kind)      # 1-based, as the reference
    d["has_children"] = take(npan, "<i4")
    d["panels_tracer"] = take(npan * nt).reshape(nt, npan).T.copy()
    if has_vort:
        d["panels_relvort"] = take(npan)
        d["panels_absvort"] = take(npan)
    assert off == len(raw), (off, len(raw))
    return d


def run_vtk_case(tmp, name, case):
    import shutil
    vtk, dump = os.path.join(tmp, name + ".vtk"), os.path.join(tmp, name + ".bin")
    subprocess.run([VTK_HARNESS, str(case), vtk, dump], check=True, cwd=tmp,
                   stdout=subprocess.DEVNULL)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **read_vtk_dump(dump))
    shutil.copy(vtk, os.path.join(HERE, name + ".vtk"))
    print(name, os.path.getsize(vtk))


def run_driver(tmp, name, binary, namelist):
    """Run a reference driver (built by make -f oracle/Makefile.ref drivers)
    in a scratch folder and copy its .dat table to tests/golden/<name>.dat."""
    import glob
    import shutil
    work = os.path.join(tmp, name)
    os.makedirs(os.path.join(work, "vtkOut"))
    shutil.copy(os.path.join(HERE, name + ".namelist"), os.path.join(work, namelist))
    subprocess.run([os.path.join(ROOT, "oracle", "_ref", binary)], check=True, cwd=work,
                   stdout=subprocess.DEVNULL, timeout=600)
    dat = glob.glob(os.path.join(work, "*.dat"))
    assert len(dat) == 1, dat
    shutil.copy(dat[0], os.path.join(HERE, name + ".dat"))
    print(name, os.path.basename(dat[0]))


def read_dump(path, forced=False):
    """Parse the stream file written by make_golden.f90."""
    raw = open(path, "rb").read()
    off = 0
    na, npart, npp, nsteps = np.frombuffer(raw, "<i4", 4, off)
    off += 16
    dt, smooth, omega = np.frombuffer(raw, "<f8", 3, off)
    off += 24

    def take(n):
        nonlocal off
        a = np.frombuffer(raw, "<f8", n, off).copy()
        off += 8 * n
        return a

    def state():
        s = {}
        s["active_x"] = take(3 * na).reshape(na, 3)
        s["active_relvort"] = take(na)
        s["active_area"] = take(na)
        s["passive_x"] = take(3 * npp).reshape(npp, 3)
        s["particles_x"] = take(3 * npart).reshape(npart, 3)
        s["particles_relvort"] = take(npart)
        s["particles_absvort"] = take(npart)
        return s

    s0, s1, sn = state(), state(), state()
    ke1 = take(na)
    ke_total = take(nsteps)
    extra = {}
    if forced:
        extra["t0"] = take(1)
        extra["active_absvort"] = take(na)
    assert off == len(raw), (off, len(raw))
    return dict(na=int(na), npart=int(npart), npp=int(npp), nsteps=int(nsteps),
                dt=float(dt), smooth=float(smooth), omega=float(omega),
                s0=s0, s1=s1, sn=sn, ke1=ke1, ke_total=ke_total, extra=extra)


def run_case(tmp, name, level, ic, nsteps, dt, smooth, t0=None):
    """Run the harness for one case; write its .npz; return the parsed dump."""
    out = os.path.join(tmp, name + ".bin")
    cmd = [HARNESS, str(level), str(ic), str(nsteps), repr(dt), repr(smooth), out]
    if t0 is not None:
        cmd.append(repr(t0))
    subprocess.run(cmd, check=True, cwd=tmp, stdout=subprocess.DEVNULL)
    d = read_dump(out, forced=t0 is not None)
    arrays = {}
    for tag, st in (("s0", d["s0"]), ("s1", d["s1"]), ("sn", d["sn"])):
        if tag == "sn" and d["nsteps"] == 1:
            continue
        for k, v in st.items():
            arrays[f"{tag}_{k}"] = v
    arrays["ke_step1"] = d["ke1"]
    arrays["total_ke"] = d["ke_total"]
    arrays.update(d["extra"])
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
    print(name, d["na"], d["npart"], d["npp"])
    return d


def main_voronoi():
    """Only the Voronoi cases; merged into an existing manifest.json."""
    if not os.path.exists(VORONOI_HARNESS):
        sys.exit("build the harness first: make -f oracle/Makefile.ref")
    cases = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (nest, ic, nsteps, dt, saved, desc) in VORONOI_CASES.items():
            d = run_voronoi_case(tmp, name, nest, ic, nsteps, dt, saved)
            cases[name] = dict(init_nest=nest, ic=ic, nsteps=nsteps, dt=d["dt"], omega=d["omega"],
                               saved_area=bool(saved), n=d["n"], description=desc)
        for name, (binary, namelist) in VORONOI_DRIVER_CASES.items():
            run_driver(tmp, name, binary, namelist)
    path = os.path.join(HERE, "manifest.json")
    with open(path) as f:
        manifest = json.load(f)
    manifest["voronoi_source"] = (
        "reference VoronoiVelocityParallelModule::BVERK4 (VoronoiVelocityParallel.f90:304-483) on "
        "a grid from VoronoiPanelsModule::New (VoronoiPanels.f90:245-315, STRIPACK), one rank; "
        "s<k>_* = state after step k (s0 initial; area = that state's tessellation), ke_step<k> / "
        "total_ke_step<k> = the stage-1 kinetic energy of step k")
    manifest["voronoi_cases"] = cases
    with open(path, "w") as f:
        json.dump(manifest, f, indent=1)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -f oracle/Makefile.ref")
    manifest, forced, advection = {}, {}, {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (level, ic, nsteps, dt, smooth, desc) in CASES.items():
            d = run_case(tmp, name, level, ic, nsteps, dt, smooth)
            manifest[name] = dict(level=level, ic=ic, nsteps=d["nsteps"], dt=d["dt"],
                                  smooth=d["smooth"], omega=d["omega"],
                                  n_active=d["na"], n_particles=d["npart"],
                                  n_passive=d["npp"], description=desc)
        for name, (level, nsteps, dt, smooth, t0, desc) in FORCED_CASES.items():
            d = run_case(tmp, name, level, 3, nsteps, dt, smooth, t0)
            forced[name] = dict(level=level, ic=3, nsteps=d["nsteps"], dt=d["dt"],
                                smooth=d["smooth"], omega=d["omega"], t0=t0,
                                n_active=d["na"], n_particles=d["npart"],
                                n_passive=d["npp"], description=desc)
        for name, (level, ic, nsteps, dt, t0, desc) in ADVECTION_CASES.items():
            d = run_case(tmp, name, level, ic, nsteps, dt, 0.05, t0)
            advection[name] = dict(level=level, ic=ic, divergent=ic == 5, nsteps=d["nsteps"],
                                   dt=d["dt"], t0=t0, n_active=d["na"], n_particles=d["npart"],
                                   n_passive=d["npp"], description=desc)
        for name, (binary, namelist) in DRIVER_CASES.items():
            run_driver(tmp, name, binary, namelist)
        for name, (case, desc) in VTK_CASES.items():
            run_vtk_case(tmp, name, case)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(dict(
            source="reference PanelVelocityParallelModule::BVERK4 (PanelVelocityParallel.f90:435-658), "
                   "grid from ParticlesEdgesPanelsModule::New (ParticlesEdgesPanels2.f90:362), "
                   "compiled with ROCm flang -O2 + MPICH, run on 1 rank",
            states="s0 = initial, s1 = after step 1, sn = after the last step (absent when nsteps == 1)",
            cases=manifest,
            forced_source="reference PanelVelocityParallelModule::StratosphereRK4 "
                          "(PanelVelocityParallel.f90:661-919) from InitStratosphere "
                          "(TracerAndVorticityDistributions.f90:793-825); extra arrays t0 and "
                          "active_absvort",
            forced_cases=forced,
            advection_source="reference AdvectionRK4 / DivergentAdvectionRK4 "
                             "(PanelVelocityParallel.f90:922-1279, winds :1415-1459); only the "
                             "positions change",
            advection_cases=advection,
            vtk_source="reference vtkOutput / vtkOutputUnif (ParticlesEdgesPanels2.f90:2116-2260) "
                       "run by make_vtk_golden.f90; <name>.vtk is the frame, <name>.npz the arrays "
                       "it was written from (vertices 1-based)",
            vtk_cases={k: v[1] for k, v in VTK_CASES.items()}), f, indent=1)


if __name__ == "__main__":
    if sys.argv[1:] == ["voronoi"]:
        main_voronoi()
    else:
        main()
        main_voronoi()
