"""The Fortran side of the boundary on a host without a GPU: the
ISO_C_BINDING host and the drop-in module report a missing device through
bve_last_error and stop with a nonzero status."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from fortran_io import write_host_input
from thesiscode_amd import grid as gridmod
from thesiscode_amd import initial_conditions as ic

HOST = os.path.join(ROOT, "thesiscode_amd", "bve_host.exe")


def has_gpu():
    from thesiscode_amd import _lib
    import ctypes
    L = _lib.load()
    h = ctypes.c_void_p()
    if L.bve_create(ctypes.byref(h), 0) == 0:
        L.bve_destroy(h)
        return True
    return False


@pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (no flang)")
def test_fortran_host_reports_missing_device(tmp_path):
    if has_gpu():
        pytest.skip("a GPU is present: the GPU tests cover this host")
    g = gridmod.cubed_sphere(1)
    z = ic.rh4_wave(g.active_x)
    inp = tmp_path / "in.bin"
    write_host_input(inp, g.active_x, z, g.active_area, g.passive_x, g.particles_x,
                     np.zeros(g.n_particles), 1, 0.01, 2 * np.pi, 0.05)
    r = subprocess.run([HOST, str(inp), str(tmp_path / "out.bin")], capture_output=True, text=True)
    assert r.returncode != 0
    assert "no HIP device" in r.stdout + r.stderr
