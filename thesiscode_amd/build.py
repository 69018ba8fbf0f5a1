"""Builds libbve.so (the HIP library, gfx950) in-tree with hipcc.

    python -m thesiscode_amd.build          # or thesiscode_amd.build.build()

The shared library lands next to this file so that it travels with the
source tree; nothing is installed.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbve.so")
SOURCES = ["bve_kernels.hip", "bve_mixed.hip", "bve_adjoint.hip", "bve_ensemble.hip", "bve_api.hip"]
# host-only C++ (no device code): compiled by the host compiler
CXX_SOURCES = ["bve_vtk.cpp"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-Wall"]
HEADERS = ["bve_common.h", "bve_kernels.h"]
ARCH = os.environ.get("BVE_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 are generated anywhere.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def cxx():
    for cand in (os.environ.get("CXX"), shutil.which("g++"), shutil.which("c++")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("no host C++ compiler found")


def build(verbose=False, extra_flags=()):
    """Compile every HIP source and link libbve.so (always rebuilds)."""
    objs = []
    build_dir = os.path.join(HERE, "_build")
    os.makedirs(build_dir, exist_ok=True)
    for src in SOURCES:
        obj = os.path.join(build_dir, src.replace(".hip", ".o"))
        cmd = [hipcc(), *HIPCC_FLAGS, *extra_flags, "-I", CSRC, "-I", os.path.join(ROOT, "include"),
               "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    for src in CXX_SOURCES:
        obj = os.path.join(build_dir, src.replace(".cpp", ".o"))
        cmd = [cxx(), *CXX_FLAGS, "-I", CSRC, "-I", os.path.join(ROOT, "include"),
               "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-pthread", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_fortran_host(verbose)
    build_copy_probe(verbose)
    return LIB


COPY_PROBE = os.path.join(HERE, "copy_probe")


def build_copy_probe(verbose=False):
    """The byte-counter calibration tool (csrc/copy_probe.hip; not part of
    libbve), run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE."""
    tmp = COPY_PROBE + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, os.path.join(CSRC, "copy_probe.hip"), "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, COPY_PROBE)
    return COPY_PROBE


FORTRAN = os.path.join(HERE, "fortran")
HOST_EXE = os.path.join(HERE, "bve_host.exe")


def flang():
    for cand in (os.environ.get("FLANG"), "/opt/rocm/llvm/bin/flang", shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    return None


def build_fortran_host(verbose=False):
    """The standalone Fortran host (fortran/bve_host.f90 over the
    ISO_C_BINDING interface), linked against libbve.so; skipped when no
    Fortran compiler is installed."""
    fc = flang()
    if fc is None:
        return None
    mod_dir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(mod_dir, exist_ok=True)
    tmp = HOST_EXE + ".tmp"
    cmd = [fc, "-O2", "-J", mod_dir, os.path.join(FORTRAN, "bve_interface.f90"),
           os.path.join(FORTRAN, "bve_host.f90"), "-L", HERE, "-lbve",
           "-Wl,-rpath,$ORIGIN", "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, HOST_EXE)
    return HOST_EXE


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
