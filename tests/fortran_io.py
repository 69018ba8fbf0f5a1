"""Input/output files of the Fortran host (thesiscode_amd/fortran/bve_host.f90)."""
import numpy as np


def write_host_input(path, ax, av, ar, px, qx, qa, n_steps, dt, omega, smooth, precision=0,
                     device=0):
    with open(path, "wb") as f:
        f.write(np.array([len(av), len(px), len(qx), n_steps, precision, device], "<i4").tobytes())
        f.write(np.array([dt, omega, smooth], "<f8").tobytes())
        for a in (ax, av, ar, px, qx, qa):
            f.write(np.ascontiguousarray(a, "<f8").tobytes())


def read_host_output(path, na, npp, npt, n_steps):
    raw = np.fromfile(path, "<f8")
    off = 0

    def take(n):
        nonlocal off
        v = raw[off:off + n]
        off += n
        return v
    out = dict(active_x=take(3 * na).reshape(na, 3), active_relvort=take(na),
               passive_x=take(3 * npp).reshape(npp, 3), particles_x=take(3 * npt).reshape(npt, 3),
               particles_relvort=take(npt), total_ke=take(n_steps))
    assert off == raw.size
    return out


def read_dat(path):
    """The reference drivers' .dat table (header line, then F24.15 columns)."""
    return np.loadtxt(path, skiprows=1)
