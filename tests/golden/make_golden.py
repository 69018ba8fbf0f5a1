"""Write the golden vectors in this folder from the reference's own Fortran.

Runs ``oracle/_ref/make_golden`` (built by ``make -f oracle/Makefile.ref``
from the reference's unmodified sources; see tests/golden/make_golden.f90)
for a fixed list of cases and stores each case's inputs and outputs as one
compressed ``.npz`` (float64 / int32 arrays only, loadable with
``allow_pickle=False``) plus ``manifest.json``.

Run from the repository root, in a container that has the reference:

    make -f oracle/Makefile.ref && python tests/golden/make_golden.py
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


def read_dump(path):
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
    assert off == len(raw), (off, len(raw))
    return dict(na=int(na), npart=int(npart), npp=int(npp), nsteps=int(nsteps),
                dt=float(dt), smooth=float(smooth), omega=float(omega),
                s0=s0, s1=s1, sn=sn, ke1=ke1, ke_total=ke_total)


def main():
    if not os.path.exists(HARNESS):
        sys.exit("build the harness first: make -f oracle/Makefile.ref")
    manifest = {}
    with tempfile.TemporaryDirectory() as tmp:
        for name, (level, ic, nsteps, dt, smooth, desc) in CASES.items():
            out = os.path.join(tmp, name + ".bin")
            subprocess.run([HARNESS, str(level), str(ic), str(nsteps), repr(dt),
                            repr(smooth), out], check=True, cwd=tmp,
                           stdout=subprocess.DEVNULL)
            d = read_dump(out)
            arrays = {}
            for tag, st in (("s0", d["s0"]), ("s1", d["s1"]), ("sn", d["sn"])):
                if tag == "sn" and d["nsteps"] == 1:
                    continue
                for k, v in st.items():
                    arrays[f"{tag}_{k}"] = v
            arrays["ke_step1"] = d["ke1"]
            arrays["total_ke"] = d["ke_total"]
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrays)
            manifest[name] = dict(level=level, ic=ic, nsteps=d["nsteps"], dt=d["dt"],
                                  smooth=d["smooth"], omega=d["omega"],
                                  n_active=d["na"], n_particles=d["npart"],
                                  n_passive=d["npp"], description=desc)
            print(name, d["na"], d["npart"], d["npp"])
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(dict(
            source="reference PanelVelocityParallelModule::BVERK4 (PanelVelocityParallel.f90:435-658), "
                   "grid from ParticlesEdgesPanelsModule::New (ParticlesEdgesPanels2.f90:362), "
                   "compiled with ROCm flang -O2 + MPICH, run on 1 rank",
            states="s0 = initial, s1 = after step 1, sn = after the last step (absent when nsteps == 1)",
            cases=manifest), f, indent=1)


if __name__ == "__main__":
    main()
